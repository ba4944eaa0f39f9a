// GPU builders of the tile plans that the generated kernels' staged forms
// read (gml_plan.cpp holds their formats and the caching):
//   - the element-tile plan of an element table: per tile of TR rows, the
//     ascending distinct vertex ids of its rows, and per row its slots'
//     tile-local indices, one byte each, four to a word;
//   - the staging plan of an uplink: per tile of grp source rows, the
//     ascending distinct destination ids of its ball rows, and for every ball
//     slot (SELL-64) the tile-local index of its entity.
// One work-group per tile: the tile's ids are sorted in LDS (bitonic), each
// sorted position gets the rank of its value among the distinct ones, the
// distinct values are written in order and every slot finds its rank by
// binary search.  A count pass gives each tile's distinct count (a scan of
// the counts places the tiles), a fill pass writes the lists and indices.
// The host builders these replace took about 1.5 s each for a 100M-tet mesh.

#include "gml_internal.h"

#include <hipcub/hipcub.hpp>

#include <climits>
#include <vector>

namespace gml {

namespace {

constexpr int kPB = 256;       // work-group size
constexpr int kPMax = 8192;    // ids of one tile at most

__device__ __forceinline__ int next_pow2(int m)
{
   int p = 1;
   while(p < m) p <<= 1;
   return p;
}

// ascending bitonic sort of s[0..np), np a power of two
__device__ void bitonic(int *s, int np)
{
   for(int k = 2; k <= np; k <<= 1)
      for(int j = k >> 1; j > 0; j >>= 1)
      {
         for(int i = threadIdx.x; i < np; i += kPB)
         {
            const int x = i ^ j;
            if(x > i)
            {
               const int a = s[i], b = s[x];
               if((a > b) == ((i & k) == 0)) { s[i] = b; s[x] = a; }
            }
         }
         __syncthreads();
      }
}

// exclusive scan of one int per thread; *total receives the sum (all threads)
__device__ int block_scan(int v, int *total)
{
   __shared__ int sc[kPB];
   sc[threadIdx.x] = v;
   __syncthreads();
   for(int o = 1; o < kPB; o <<= 1)
   {
      const int t = threadIdx.x >= (unsigned)o ? sc[threadIdx.x - o] : 0;
      __syncthreads();
      sc[threadIdx.x] += t;
      __syncthreads();
   }
   const int incl = sc[threadIdx.x];
   *total = sc[kPB - 1];
   __syncthreads();
   return incl - v;
}

// rk[i] (i < m, s sorted): the rank of s[i] among the distinct values of
// s[0..m); returns the number of distinct values
__device__ int ranks(const int *s, int m, unsigned short *rk)
{
   const int per = (m + kPB - 1) / kPB;
   const int i0 = min(m, (int)threadIdx.x * per), i1 = min(m, i0 + per);
   int c = 0;
   for(int i = i0; i < i1; i++) c += (i == 0 || s[i] != s[i - 1]);
   int nd = 0;
   int r = block_scan(c, &nd);
   for(int i = i0; i < i1; i++)
   {
      r += (i == 0 || s[i] != s[i - 1]);
      rk[i] = (unsigned short)(r - 1);
   }
   __syncthreads();
   return nd;
}

// first position of v in the sorted s[0..m) (v is present)
__device__ __forceinline__ int lower(const int *s, int m, int v)
{
   int lo = 0, hi = m;
   while(lo < hi)
   {
      const int mid = (lo + hi) >> 1;
      if(s[mid] < v) lo = mid + 1; else hi = mid;
   }
   return lo;
}

// --- element tiles ----------------------------------------------------------
// the tile's ids (ele rows [a, b), w each) into s, padded to a power of two;
// flags |= 1 for an id outside [0, nv)
__device__ int load_rows(const int *ele, long long a, long long b, int w, long long nv, int *s, int *flags)
{
   const int m = (int)((b - a) * w), np = next_pow2(max(m, 1));
   bool bad = false;
   for(int i = threadIdx.x; i < np; i += kPB)
   {
      int v = INT_MAX;
      if(i < m)
      {
         v = ele[a * w + i];
         if(v < 0 || (long long)v >= nv) bad = true;
      }
      s[i] = v;
   }
   if(bad) atomicOr(flags, 1);
   __syncthreads();
   bitonic(s, np);
   return m;
}

__global__ void __launch_bounds__(kPB) k_tplan_count(const int *__restrict__ ele, long long n, int w, long long nv, int TR,
                                                     unsigned int *__restrict__ cnt, int *flags)
{
   __shared__ int s[kPMax];
   __shared__ unsigned short rk[kPMax];
   const long long a = (long long)blockIdx.x * TR, b = min(n, a + TR);
   const int m = load_rows(ele, a, b, w, nv, s, flags);
   const int nd = ranks(s, m, rk);
   if(threadIdx.x == 0)
   {
      cnt[blockIdx.x] = (unsigned int)nd;
      if(nd > 256) atomicOr(flags, 2);
   }
}

__global__ void __launch_bounds__(kPB) k_tplan_fill(const int *__restrict__ ele, long long n, int w, int wq, long long nv,
                                                    int TR, const unsigned int *__restrict__ tofs, int *__restrict__ tlist,
                                                    unsigned int *__restrict__ lid, int *flags)
{
   __shared__ int s[kPMax];
   __shared__ unsigned short rk[kPMax];
   const long long a = (long long)blockIdx.x * TR, b = min(n, a + TR);
   const int m = load_rows(ele, a, b, w, nv, s, flags);
   ranks(s, m, rk);
   const unsigned int o = tofs[blockIdx.x];
   for(int i = threadIdx.x; i < m; i += kPB)
      if(i == 0 || s[i] != s[i - 1]) tlist[o + rk[i]] = s[i];
   for(long long r = a + threadIdx.x; r < b; r += kPB)
      for(int q = 0; q < wq; q++)
      {
         unsigned int word = 0;
         for(int j = 4 * q; j < min(w, 4 * q + 4); j++)
            word |= (unsigned int)rk[lower(s, m, ele[r * w + j])] << (8 * (j & 3));
         lid[r * wq + q] = word;
      }
}

// --- uplink staging tiles ---------------------------------------------------
// the tile's ball entries (rows [r0, r1), deg[r] entries each, SELL-64) as
// destination ids into s, padded to a power of two; flags |= 4 if more
// than kPMax
__device__ int load_ball(const int *ent, const unsigned int *off, const int *deg, long long r0, long long r1, int *s,
                         int *rofs, int *flags)
{
   const int nr = (int)(r1 - r0);
   const int d = (int)threadIdx.x < nr ? deg[r0 + threadIdx.x] : 0;
   int m = 0;
   rofs[threadIdx.x] = block_scan(d, &m);
   __syncthreads();
   if(m > kPMax)
   {
      if(threadIdx.x == 0) atomicOr(flags, 4);
      return -1;
   }
   if((int)threadIdx.x < nr)
   {
      const long long r = r0 + threadIdx.x;
      const size_t base = (size_t)off[r >> 6] * 64 + (r & 63);
      for(int k = 0; k < d; k++) s[rofs[threadIdx.x] + k] = ent[base + (size_t)k * 64] >> 4;
   }
   const int np = next_pow2(max(m, 1));
   for(int i = m + threadIdx.x; i < np; i += kPB) s[i] = INT_MAX;
   __syncthreads();
   bitonic(s, np);
   return m;
}

__global__ void __launch_bounds__(kPB) k_splan_count(const int *__restrict__ ent, const unsigned int *__restrict__ off,
                                                     const int *__restrict__ deg, long long nrows, int grp,
                                                     unsigned int *__restrict__ cnt, int *flags)
{
   __shared__ int s[kPMax];
   __shared__ unsigned short rk[kPMax];
   __shared__ int rofs[kPB];
   const long long r0 = (long long)blockIdx.x * grp, r1 = min(nrows, r0 + grp);
   const int m = load_ball(ent, off, deg, r0, r1, s, rofs, flags);
   if(m < 0) return;
   const int nd = ranks(s, m, rk);
   if(threadIdx.x == 0) cnt[blockIdx.x] = (unsigned int)nd;
}

__global__ void __launch_bounds__(kPB) k_splan_fill(const int *__restrict__ ent, const unsigned int *__restrict__ off,
                                                    const int *__restrict__ deg, long long nrows, int grp,
                                                    const unsigned int *__restrict__ tofs, int *__restrict__ tlist,
                                                    unsigned short *__restrict__ lidx, int *flags)
{
   __shared__ int s[kPMax];
   __shared__ unsigned short rk[kPMax];
   __shared__ int rofs[kPB];
   const long long r0 = (long long)blockIdx.x * grp, r1 = min(nrows, r0 + grp);
   const int m = load_ball(ent, off, deg, r0, r1, s, rofs, flags);
   if(m < 0) return;
   ranks(s, m, rk);
   const unsigned int o = tofs[blockIdx.x];
   for(int i = threadIdx.x; i < m; i += kPB)
      if(i == 0 || s[i] != s[i - 1]) tlist[o + rk[i]] = s[i];
   if((long long)threadIdx.x < r1 - r0)
   {
      const long long r = r0 + threadIdx.x;
      const size_t base = (size_t)off[r >> 6] * 64 + (r & 63);
      const int d = deg[r];
      for(int k = 0; k < d; k++)
      {
         const size_t sl = base + (size_t)k * 64;
         lidx[sl] = rk[lower(s, m, ent[sl] >> 4)];
      }
   }
}

bool ok_(hipError_t e) { return e == hipSuccess; }

// tile offsets from the per-tile counts cnt[0..ntile) (cnt[ntile] = 0): the
// exclusive scan into tofs[0..ntile], and the largest count
bool ScanCounts(GmlCtx *g, unsigned int *cnt, long long ntile, unsigned int *tofs, long long *sum, int *maxn)
{
   hipStream_t s = g->stream;
   size_t t1 = 0, t2 = 0;
   unsigned int *mx = (unsigned int *)DevAlloc(g, 16);
   bool ok = mx && ok_(hipcub::DeviceScan::ExclusiveSum(nullptr, t1, cnt, tofs, (int)(ntile + 1), s))
          && ok_(hipcub::DeviceReduce::Max(nullptr, t2, cnt, mx, (int)ntile, s));
   void *tmp = ok ? DevAlloc(g, std::max(t1, t2)) : nullptr;
   ok = ok && tmp;
   size_t t = std::max(t1, t2);
   ok = ok && ok_(hipcub::DeviceScan::ExclusiveSum(tmp, t, cnt, tofs, (int)(ntile + 1), s));
   t = std::max(t1, t2);
   ok = ok && ok_(hipcub::DeviceReduce::Max(tmp, t, cnt, mx, (int)ntile, s));
   unsigned int h[2] = {0, 0};
   ok = ok && DevCopy(g, &h[0], tofs + ntile, 4, 2) && DevCopy(g, &h[1], mx, 4, 2);
   DevFree(g, tmp);
   DevFree(g, mx);
   *sum = h[0];
   *maxn = (int)h[1];
   return ok;
}

} // namespace

// Element-tile plan on the GPU.  Returns 1 with the arrays allocated, 0 if a
// tile of TR rows reaches more than 256 vertices (try smaller tiles), -1 if
// an id is outside the vertex table or the GPU work failed.
int GpuTilePlan(GmlCtx *g, const int *ele, long long n, int w, int wq, long long nv, int TR, unsigned int **tofs,
                int **tlist, unsigned int **lid, long long *sum, int *maxn)
{
   *tofs = nullptr; *tlist = nullptr; *lid = nullptr;
   if(n < 1 || w < 1 || w > 16 || (long long)TR * w > kPMax) return -1;
   const long long ntile = (n + TR - 1) / TR;
   hipStream_t s = g->stream;
   unsigned int *cnt = (unsigned int *)DevAlloc(g, (size_t)(ntile + 1) * 4);
   int *flags = (int *)DevAlloc(g, 4);
   *tofs = (unsigned int *)DevAlloc(g, (size_t)(ntile + 1) * 4);
   int f = 0;
   bool ok = cnt && flags && *tofs && ok_(hipMemsetAsync(cnt, 0, (size_t)(ntile + 1) * 4, s))
          && ok_(hipMemsetAsync(flags, 0, 4, s));
   if(ok)
   {
      hipLaunchKernelGGL(k_tplan_count, dim3((unsigned int)ntile), dim3(kPB), 0, s, ele, n, w, nv, TR, cnt, flags);
      ok = ok_(hipGetLastError()) && DevCopy(g, &f, flags, 4, 2);
   }
   int rc = ok ? (f & 1 ? -1 : (f & 2 ? 0 : 1)) : -1;
   if(rc == 1)
   {
      rc = ScanCounts(g, cnt, ntile, *tofs, sum, maxn) ? 1 : -1;
      if(rc == 1)
      {
         *tlist = (int *)DevAlloc(g, (size_t)*sum * 4 + 4);
         *lid = (unsigned int *)DevAlloc(g, (size_t)n * wq * 4);
         if(*tlist && *lid)
         {
            hipLaunchKernelGGL(k_tplan_fill, dim3((unsigned int)ntile), dim3(kPB), 0, s, ele, n, w, wq, nv, TR, *tofs,
                               *tlist, *lid, flags);
            rc = ok_(hipGetLastError()) && ok_(hipStreamSynchronize(s)) ? 1 : -1;
         }
         else
            rc = -1;
      }
   }
   DevFree(g, cnt);
   DevFree(g, flags);
   if(rc != 1)
   {
      DevFree(g, *tofs); DevFree(g, *tlist); DevFree(g, *lid);
      *tofs = nullptr; *tlist = nullptr; *lid = nullptr;
   }
   return rc;
}

// Uplink staging plan on the GPU (rows of at most 8192 / grp entries each,
// which the base ball rows of the staged kernels are).  lidx must hold one
// entry per ball slot; slots past a row's degree are set to 0.  Returns 1
// with tofs / tlist allocated, or -1.
int GpuStagePlan(GmlCtx *g, const int *ent, const unsigned int *off, const int *deg, long long nrows, int grp,
                 size_t nslot, unsigned int **tofs, int **tlist, unsigned short *lidx, long long *sum, int *maxn)
{
   *tofs = nullptr; *tlist = nullptr;
   if(nrows < 1 || grp < 1 || grp > kPB) return -1;
   const long long ntile = (nrows + grp - 1) / grp;
   hipStream_t s = g->stream;
   unsigned int *cnt = (unsigned int *)DevAlloc(g, (size_t)(ntile + 1) * 4);
   int *flags = (int *)DevAlloc(g, 4);
   *tofs = (unsigned int *)DevAlloc(g, (size_t)(ntile + 1) * 4);
   int f = 0;
   bool ok = cnt && flags && *tofs && ok_(hipMemsetAsync(cnt, 0, (size_t)(ntile + 1) * 4, s))
          && ok_(hipMemsetAsync(flags, 0, 4, s)) && ok_(hipMemsetAsync(lidx, 0, nslot * 2, s));
   if(ok)
   {
      hipLaunchKernelGGL(k_splan_count, dim3((unsigned int)ntile), dim3(kPB), 0, s, ent, off, deg, nrows, grp, cnt, flags);
      ok = ok_(hipGetLastError()) && DevCopy(g, &f, flags, 4, 2) && f == 0;
   }
   ok = ok && ScanCounts(g, cnt, ntile, *tofs, sum, maxn) && *maxn <= 65535;
   if(ok)
   {
      *tlist = (int *)DevAlloc(g, (size_t)*sum * 4 + 4);
      ok = *tlist != nullptr;
   }
   if(ok)
   {
      hipLaunchKernelGGL(k_splan_fill, dim3((unsigned int)ntile), dim3(kPB), 0, s, ent, off, deg, nrows, grp, *tofs,
                         *tlist, lidx, flags);
      ok = ok_(hipGetLastError()) && ok_(hipStreamSynchronize(s));
   }
   DevFree(g, cnt);
   DevFree(g, flags);
   if(!ok)
   {
      DevFree(g, *tofs); DevFree(g, *tlist);
      *tofs = nullptr; *tlist = nullptr;
      return -1;
   }
   return 1;
}

} // namespace gml
