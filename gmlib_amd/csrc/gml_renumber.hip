// Hilbert renumbering of a mesh on the GPU: vertices sorted along a 3D
// Hilbert curve of their bounding box, elements along the same curve by
// their barycentre, element tables renumbered.  Every kernel of the engine
// assumes such an order (Hilbert-ordered, locality-preserving numbering);
// the reference takes it from an external library (LPlib) before GMlib is
// called (README.md:50, Documentation/GMlib3_en.tex:340).
//
// Keys: 21 bits per axis (63-bit Hilbert index, gml_hilbert.h); sorting is
// a stable radix sort of (key, index) pairs, so equal keys keep the input
// order and the result is deterministic.  The optional degree class sets the
// vertex key's top bit for vertices in more than BasDeg elements, which puts
// them after all the others (the "-gmlib" preprocessing the reference's
// documentation asks for: low-degree vertices first, high-degree ones at the
// end, Documentation/GMlib3_en.tex:401).

#include "gml_internal.h"
#include "gml_device.h"
#include "gml_hilbert.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <thread>

namespace gml {

static const int kHilBits = 21;

__device__ __forceinline__ unsigned Quant(float x, float lo, float sc)
{
   float q = (x - lo) * sc;
   q = fminf(fmaxf(q, 0.f), (float)((1u << kHilBits) - 1));
   return (unsigned)q;
}

// element degree of every vertex (deg zeroed before)
__global__ void k_vertex_degree(long long ne, int nnod, const int *__restrict__ ele, int *__restrict__ deg)
{
   long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(i < ne * nnod) atomicAdd(&deg[ele[i]], 1);
}

// Hilbert key of each vertex; with deg, bit 63 set for degrees above basdeg
// (the key uses 63 bits, so the high class sorts after the low one)
__global__ void k_vertex_keys(long long nv, const v4f *__restrict__ crd, v4f lo, v4f sc, const int *__restrict__ deg,
                              int basdeg, unsigned long long *__restrict__ key, int *__restrict__ idx)
{
   long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(v >= nv) return;
   v4f p = crd[v];
   unsigned long long k = HilbertKey(Quant(p.x, lo.x, sc.x), Quant(p.y, lo.y, sc.y), Quant(p.z, lo.z, sc.z), kHilBits);
   if(deg && deg[v] > basdeg) k |= 1ull << 63;
   key[v] = k;
   idx[v] = (int)v;
}

__global__ void k_element_keys(long long ne, int nnod, const int *__restrict__ ele, const v4f *__restrict__ crd,
                               v4f lo, v4f sc, unsigned long long *__restrict__ key, int *__restrict__ idx)
{
   long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(e >= ne) return;
   v4f c = {0.f, 0.f, 0.f, 0.f};
   for(int j = 0; j < nnod; j++) c += crd[ele[e * nnod + j]];
   c *= 1.f / (float)nnod;
   key[e] = HilbertKey(Quant(c.x, lo.x, sc.x), Quant(c.y, lo.y, sc.y), Quant(c.z, lo.z, sc.z), kHilBits);
   idx[e] = (int)e;
}

// inv[old[i]] = i
__global__ void k_invert(long long n, const int *__restrict__ old, int *__restrict__ inv)
{
   long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(i < n) inv[old[i]] = (int)i;
}

__global__ void k_gather_vertices(long long nv, const int *__restrict__ old, const v4f *__restrict__ crd,
                                  const int *__restrict__ ref, v4f *__restrict__ crd2, int *__restrict__ ref2)
{
   long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(i >= nv) return;
   const int o = old[i];
   crd2[i] = crd[o];
   if(ref) ref2[i] = ref[o];
}

__global__ void k_gather_elements(long long ne, int nnod, const int *__restrict__ old, const int *__restrict__ inv,
                                  const int *__restrict__ ele, const int *__restrict__ ref, int *__restrict__ ele2,
                                  int *__restrict__ ref2)
{
   long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(i >= ne) return;
   const int o = old[i];
   for(int j = 0; j < nnod; j++) ele2[i * nnod + j] = inv[ele[(long long)o * nnod + j]];
   if(ref) ref2[i] = ref[o];
}

} // namespace gml

using namespace gml;

extern "C" {

int GmlxHilbertRenumber(int NmbVer, float *crd, int *VerRef, int NmbEle, int nnod, int *ele, int *EleRef,
                        int *VerOld, int *EleOld, int BasDeg)
{
   if(NmbVer < 1 || NmbEle < 0 || nnod < 1 || nnod > 8 || !crd || (NmbEle && !ele) || BasDeg < 0) return 0;
   const long long nv = NmbVer, ne = NmbEle;
   for(long long i = 0; i < ne * nnod; i++)
      if(ele[i] < 0 || ele[i] >= NmbVer) return 0;
   // bounding box of the vertices (host: the data are there already)
   float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
   for(long long v = 0; v < nv; v++)
      for(int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], crd[v * 4 + a]); hi[a] = std::max(hi[a], crd[v * 4 + a]); }
   const float top = (float)((1u << kHilBits) - 1);
   v4f vlo = {lo[0], lo[1], lo[2], 0.f}, vsc;
   for(int a = 0; a < 3; a++) vsc[a] = hi[a] > lo[a] ? top / (hi[a] - lo[a]) : 0.f;
   vsc[3] = 0.f;

   hipStream_t s = nullptr;
   if(hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return 0;
   const long long nmax = std::max(nv, ne);
   v4f *dc = nullptr, *dc2 = nullptr;
   int *de = nullptr, *de2 = nullptr, *dvr = nullptr, *dvr2 = nullptr, *der = nullptr, *der2 = nullptr;
   int *idx = nullptr, *idx2 = nullptr, *vold = nullptr, *vinv = nullptr, *eold = nullptr;
   unsigned long long *key = nullptr, *key2 = nullptr;
   void *tmp = nullptr;
   size_t tb = 0, tb2 = 0;
   int rc = 0;
   const unsigned B = 256;
   auto blocks = [&](long long n) { return dim3((unsigned)((n + B - 1) / B)); };
   bool ok = hipMalloc(&dc, nv * 16) == hipSuccess && hipMalloc(&dc2, nv * 16) == hipSuccess
          && hipMalloc(&key, nmax * 8) == hipSuccess && hipMalloc(&key2, nmax * 8) == hipSuccess
          && hipMalloc(&idx, nmax * 4) == hipSuccess && hipMalloc(&idx2, nmax * 4) == hipSuccess
          && hipMalloc(&vold, nv * 4) == hipSuccess && hipMalloc(&vinv, nv * 4) == hipSuccess
          && (!VerRef || (hipMalloc(&dvr, nv * 4) == hipSuccess && hipMalloc(&dvr2, nv * 4) == hipSuccess))
          && (!ne || (hipMalloc(&de, ne * nnod * 4) == hipSuccess && hipMalloc(&de2, ne * nnod * 4) == hipSuccess
                      && hipMalloc(&eold, ne * 4) == hipSuccess))
          && (!ne || !EleRef || (hipMalloc(&der, ne * 4) == hipSuccess && hipMalloc(&der2, ne * 4) == hipSuccess));
   if(ok)
      ok = hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key, key2, idx, idx2, (int)nv, 0, 64, s) == hipSuccess
        && (!ne || hipcub::DeviceRadixSort::SortPairs(nullptr, tb2, key, key2, idx, idx2, (int)ne, 0, 3 * kHilBits, s)
                   == hipSuccess)
        && hipMalloc(&tmp, std::max(tb, tb2) + 16) == hipSuccess;
   if(ok)
   {
      tb = std::max(tb, tb2) + 16;
      ok = hipMemcpyAsync(dc, crd, nv * 16, hipMemcpyHostToDevice, s) == hipSuccess
        && (!VerRef || hipMemcpyAsync(dvr, VerRef, nv * 4, hipMemcpyHostToDevice, s) == hipSuccess)
        && (!ne || hipMemcpyAsync(de, ele, ne * nnod * 4, hipMemcpyHostToDevice, s) == hipSuccess)
        && (!ne || !EleRef || hipMemcpyAsync(der, EleRef, ne * 4, hipMemcpyHostToDevice, s) == hipSuccess);
   }
   if(ok)
   {
      // vertices: keys, sort, new -> old and old -> new, permuted coordinates
      // (with a degree class: the degrees counted in vinv, rewritten below)
      if(BasDeg > 0)
      {
         ok = hipMemsetAsync(vinv, 0, nv * 4, s) == hipSuccess;
         if(ne) hipLaunchKernelGGL(k_vertex_degree, blocks(ne * nnod), dim3(B), 0, s, ne, nnod, de, vinv);
      }
      hipLaunchKernelGGL(k_vertex_keys, blocks(nv), dim3(B), 0, s, nv, dc, vlo, vsc, BasDeg > 0 ? vinv : nullptr,
                         BasDeg, key, idx);
      size_t t = tb;
      ok = ok && hipcub::DeviceRadixSort::SortPairs(tmp, t, key, key2, idx, vold, (int)nv, 0,
                                                    BasDeg > 0 ? 64 : 3 * kHilBits, s) == hipSuccess;
      hipLaunchKernelGGL(k_invert, blocks(nv), dim3(B), 0, s, nv, vold, vinv);
      hipLaunchKernelGGL(k_gather_vertices, blocks(nv), dim3(B), 0, s, nv, vold, dc, dvr, dc2, dvr2);
      // elements: barycentre keys (original coordinates), sort, renumbered tables
      if(ok && ne)
      {
         hipLaunchKernelGGL(k_element_keys, blocks(ne), dim3(B), 0, s, ne, nnod, de, dc, vlo, vsc, key, idx);
         t = tb;
         ok = hipcub::DeviceRadixSort::SortPairs(tmp, t, key, key2, idx, eold, (int)ne, 0, 3 * kHilBits, s) == hipSuccess;
         hipLaunchKernelGGL(k_gather_elements, blocks(ne), dim3(B), 0, s, ne, nnod, eold, vinv, de, der, de2, der2);
      }
      ok = ok && hipGetLastError() == hipSuccess
        && hipMemcpyAsync(crd, dc2, nv * 16, hipMemcpyDeviceToHost, s) == hipSuccess
        && (!VerRef || hipMemcpyAsync(VerRef, dvr2, nv * 4, hipMemcpyDeviceToHost, s) == hipSuccess)
        && (!VerOld || hipMemcpyAsync(VerOld, vold, nv * 4, hipMemcpyDeviceToHost, s) == hipSuccess)
        && (!ne || hipMemcpyAsync(ele, de2, ne * nnod * 4, hipMemcpyDeviceToHost, s) == hipSuccess)
        && (!ne || !EleRef || hipMemcpyAsync(EleRef, der2, ne * 4, hipMemcpyDeviceToHost, s) == hipSuccess)
        && (!ne || !EleOld || hipMemcpyAsync(EleOld, eold, ne * 4, hipMemcpyDeviceToHost, s) == hipSuccess)
        && hipStreamSynchronize(s) == hipSuccess;
      rc = ok ? 1 : 0;
   }
   (void)hipGetLastError();
   for(void *p : {(void *)dc, (void *)dc2, (void *)de, (void *)de2, (void *)dvr, (void *)dvr2, (void *)der, (void *)der2,
                  (void *)idx, (void *)idx2, (void *)vold, (void *)vinv, (void *)eold, (void *)key, (void *)key2, tmp})
      if(p) (void)hipFree(p);
   (void)hipStreamDestroy(s);
   return rc;
}

} // extern "C"
