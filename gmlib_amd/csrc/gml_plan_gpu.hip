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

// --- operator-form sweep plan ------------------------------------------------
// Per tile of TV vertices (TV <= 256): the halo (the vertices of the tile's
// ball tets outside the tile, ascending) and, for every tile vertex, its
// 1-ring as 16-bit entries (tile-local index << 4 | count), ascending local
// index, counts above 15 split, the vertex itself only with the count beyond
// its degree.  Dynamic LDS: the tile's tets, later its unique halo (kOpT
// ints), then the halo candidates, later each thread's ring scratch (kOpH
// ints).  A tile that needs more (a degree above kOpDeg, more tets or halo
// candidates) sets flag 1 and the host builds the plan; a local index above
// 4095 or a ring longer than 65535 entries sets flag 2 (no plan).
constexpr int kOpT = 8192, kOpH = 16384, kOpDeg = 32;
constexpr size_t kOpLds = (size_t)(kOpT + kOpH) * 4;

// sum over i < m of f(i), exclusive prefix for each thread's chunk of
// ceil(m / kPB) consecutive i; *total receives the whole sum
template<class F>
__device__ int chunk_scan(int m, F f, int *total)
{
   const int per = (m + kPB - 1) / kPB;
   const int i0 = min(m, (int)threadIdx.x * per), i1 = min(m, i0 + per);
   int c = 0;
   for(int i = i0; i < i1; i++) c += f(i);
   return block_scan(c, total);
}

// the tile's unique halo into U (ascending), its count returned (-1: the
// tile takes the host path)
__device__ int op_halo(const int *ent, const unsigned int *off, const int *deg, const int4 *tet, long long v0,
                       long long v1, int *T, int *H, int *flags)
{
   __shared__ int rofs[kPB];
   const int nr = (int)(v1 - v0);
   const int d = (int)threadIdx.x < nr ? deg[v0 + threadIdx.x] : 0;
   __shared__ int big;
   if(threadIdx.x == 0) big = 0;
   __syncthreads();
   if(d > kOpDeg) big = 1;
   int m = 0;
   rofs[threadIdx.x] = block_scan(d, &m);
   __syncthreads();
   if(big || m > kOpT) { if(threadIdx.x == 0) atomicOr(flags, 1); return -1; }
   if((int)threadIdx.x < nr)
   {
      const long long v = v0 + threadIdx.x;
      const size_t base = (size_t)off[v >> 6] * 64 + (v & 63);
      for(int k = 0; k < d; k++) T[rofs[threadIdx.x] + k] = ent[base + (size_t)k * 64] >> 4;
   }
   int np = next_pow2(max(m, 1));
   for(int i = m + threadIdx.x; i < np; i += kPB) T[i] = INT_MAX;
   __syncthreads();
   bitonic(T, np);
   // the outside vertices of each distinct tet, placed by a scan
   auto outside = [&](int i) -> int {
      if(i > 0 && T[i] == T[i - 1]) return 0;
      const int4 q = tet[T[i]];
      return (q.x < v0 || q.x >= v1) + (q.y < v0 || q.y >= v1) + (q.z < v0 || q.z >= v1) + (q.w < v0 || q.w >= v1);
   };
   int nc = 0;
   int o = chunk_scan(m, outside, &nc);
   if(nc > kOpH) { if(threadIdx.x == 0) atomicOr(flags, 1); return -1; }
   {
      const int per = (m + kPB - 1) / kPB;
      const int i0 = min(m, (int)threadIdx.x * per), i1 = min(m, i0 + per);
      for(int i = i0; i < i1; i++)
      {
         if(i > 0 && T[i] == T[i - 1]) continue;
         const int4 q = tet[T[i]];
         const int x[4] = {q.x, q.y, q.z, q.w};
         for(int j = 0; j < 4; j++)
            if(x[j] < v0 || x[j] >= v1) H[o++] = x[j];
      }
   }
   np = next_pow2(max(nc, 1));
   for(int i = nc + threadIdx.x; i < np; i += kPB) H[i] = INT_MAX;
   __syncthreads();
   bitonic(H, np);
   // unique halo, compacted over the tets (no longer needed)
   int nh = 0;
   o = chunk_scan(nc, [&](int i) -> int { return i == 0 || H[i] != H[i - 1]; }, &nh);
   if(nh > kOpT) { if(threadIdx.x == 0) atomicOr(flags, 1); return -1; }
   {
      const int per = (nc + kPB - 1) / kPB;
      const int i0 = min(nc, (int)threadIdx.x * per), i1 = min(nc, i0 + per);
      for(int i = i0; i < i1; i++)
         if(i == 0 || H[i] != H[i - 1]) T[o++] = H[i];
   }
   __syncthreads();
   return nh;
}

// the ring of tile vertex v0 + threadIdx.x: its entries into out[k * 64]
// (fill) or only counted; returns the entry count (-1: an index over 4095)
__device__ int op_ring(const int *ent, const unsigned int *off, const int *deg, const int4 *tet, long long v0,
                       long long v1, int TV, const int *U, int nh, unsigned short *R, unsigned short *out)
{
   const long long v = v0 + threadIdx.x;
   const size_t base = (size_t)off[v >> 6] * 64 + (v & 63);
   const int d = deg[v];
   int n = 0;
   for(int k = 0; k < d; k++)
   {
      const int4 q = tet[ent[base + (size_t)k * 64] >> 4];
      const int x[4] = {q.x, q.y, q.z, q.w};
      for(int j = 0; j < 4; j++)
      {
         const int l = (x[j] >= v0 && x[j] < v1) ? (int)(x[j] - v0) : TV + lower(U, nh, x[j]);
         if(l > 4095) return -1;
         // insertion into the ascending scratch
         int p = n++;
         while(p > 0 && R[(p - 1) * kPB] > l) { R[p * kPB] = R[(p - 1) * kPB]; p--; }
         R[p * kPB] = (unsigned short)l;
      }
   }
   const int self = (int)(v - v0);
   int e = 0;
   for(int i = 0; i < n;)
   {
      int j = i;
      while(j < n && R[j * kPB] == R[i * kPB]) j++;
      int c = j - i;
      if(R[i * kPB] == self) c = c > d ? c - d : 0;
      for(; c > 0; c -= min(c, 15))
      {
         if(out) out[(size_t)e * 64] = (unsigned short)((R[i * kPB] << 4) | min(c, 15));
         e++;
      }
      i = j;
   }
   return e;
}

__global__ void __launch_bounds__(kPB) k_opplan_count(const int *__restrict__ ent, const unsigned int *__restrict__ off,
                                                      const int *__restrict__ deg, const int4 *__restrict__ tet,
                                                      long long nv, int TV, unsigned int *__restrict__ hcnt,
                                                      unsigned int *__restrict__ rlen, int *flags)
{
   extern __shared__ int lds[];
   int *T = lds, *H = lds + kOpT;
   const long long v0 = (long long)blockIdx.x * TV, v1 = min(nv, v0 + TV);
   const int nh = op_halo(ent, off, deg, tet, v0, v1, T, H, flags);
   if(nh < 0) return;
   if(threadIdx.x == 0) hcnt[blockIdx.x] = (unsigned int)nh;
   if((long long)threadIdx.x < v1 - v0)
   {
      unsigned short *R = reinterpret_cast<unsigned short *>(H) + threadIdx.x;
      const int e = op_ring(ent, off, deg, tet, v0, v1, TV, T, nh, R, nullptr);
      if(e < 0 || e > 65535) atomicOr(flags, 2);
      rlen[v0 + threadIdx.x] = (unsigned int)max(e, 0);
   }
}

__global__ void __launch_bounds__(kPB) k_opplan_fill(const int *__restrict__ ent, const unsigned int *__restrict__ off,
                                                     const int *__restrict__ deg, const int4 *__restrict__ tet,
                                                     long long nv, int TV, const unsigned int *__restrict__ vofs,
                                                     int *__restrict__ vlist, const unsigned int *__restrict__ noff,
                                                     unsigned short *__restrict__ nbr, int *flags)
{
   extern __shared__ int lds[];
   int *T = lds, *H = lds + kOpT;
   const long long v0 = (long long)blockIdx.x * TV, v1 = min(nv, v0 + TV);
   const int nh = op_halo(ent, off, deg, tet, v0, v1, T, H, flags);
   if(nh < 0) return;
   const unsigned int o = vofs[blockIdx.x];
   for(int i = threadIdx.x; i < nh; i += kPB) vlist[o + i] = T[i];
   if((long long)threadIdx.x < v1 - v0)
   {
      const long long v = v0 + threadIdx.x;
      unsigned short *R = reinterpret_cast<unsigned short *>(H) + threadIdx.x;
      op_ring(ent, off, deg, tet, v0, v1, TV, T, nh, R, nbr + (size_t)noff[v >> 6] * 64 + (v & 63));
   }
}

// per 64-vertex slab: the longest ring (wcnt[c]) and the ring total (ssum[c])
__global__ void k_opplan_slabs(const unsigned int *__restrict__ rlen, long long nv, long long nchunk,
                               unsigned int *__restrict__ wcnt, unsigned int *__restrict__ ssum)
{
   const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(c >= nchunk) return;
   unsigned int w = 0, t = 0;
   for(long long v = c * 64; v < min(nv, c * 64 + 64); v++) { w = max(w, rlen[v]); t += rlen[v]; }
   wcnt[c] = w;
   ssum[c] = t;
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

// Operator-form sweep plan on the GPU (tiles of TV <= 256 vertices).
// Returns 1 with vofs / vlist / noff / nbr allocated and the counts set, 0 if
// a tile exceeds the GPU builder's limits (the host builds the plan), -1 if
// the plan cannot exist (a tile-local index over 4095, a ring over 65535
// entries) or the GPU work failed.
int GpuSweepOpPlan(GmlCtx *g, const int *ent, const unsigned int *off, const int *deg, const int *tet, long long nv,
                   int TV, unsigned int **vofs, int **vlist, unsigned int **noff, unsigned short **nbr, long long *nslab,
                   long long *sum_halo, int *max_halo, long long *sum_ring)
{
   *vofs = nullptr; *vlist = nullptr; *noff = nullptr; *nbr = nullptr;
   if(nv < 1 || TV < 64 || TV > kPB || (TV & 63)) return 0;
   static bool attr = false;
   if(!attr)
   {
      if(hipFuncSetAttribute((const void *)k_opplan_count, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kOpLds) != hipSuccess
      || hipFuncSetAttribute((const void *)k_opplan_fill, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kOpLds) != hipSuccess)
      {
         (void)hipGetLastError();
         return 0;
      }
      attr = true;
   }
   const long long ntile = (nv + TV - 1) / TV, nchunk = (nv + 63) / 64;
   hipStream_t s = g->stream;
   unsigned int *hcnt = (unsigned int *)DevAlloc(g, (size_t)(ntile + 1) * 4);
   unsigned int *rlen = (unsigned int *)DevAlloc(g, (size_t)nv * 4);
   unsigned int *wcnt = (unsigned int *)DevAlloc(g, (size_t)(nchunk + 1) * 4);
   unsigned int *ssum = (unsigned int *)DevAlloc(g, (size_t)(nchunk + 1) * 4);
   unsigned int *soff = (unsigned int *)DevAlloc(g, (size_t)(nchunk + 1) * 4);
   int *flags = (int *)DevAlloc(g, 4);
   *vofs = (unsigned int *)DevAlloc(g, (size_t)(ntile + 1) * 4);
   *noff = (unsigned int *)DevAlloc(g, (size_t)(nchunk + 1) * 4);
   int f = 0, rc = -1;
   int wmax = 0;
   long long rs = 0;
   bool ok = hcnt && rlen && wcnt && ssum && soff && flags && *vofs && *noff
          && ok_(hipMemsetAsync(hcnt, 0, (size_t)(ntile + 1) * 4, s)) && ok_(hipMemsetAsync(wcnt, 0, (size_t)(nchunk + 1) * 4, s))
          && ok_(hipMemsetAsync(ssum, 0, (size_t)(nchunk + 1) * 4, s)) && ok_(hipMemsetAsync(flags, 0, 4, s));
   if(ok)
   {
      hipLaunchKernelGGL(k_opplan_count, dim3((unsigned int)ntile), dim3(kPB), kOpLds, s, ent, off, deg, (const int4 *)tet,
                         nv, TV, hcnt, rlen, flags);
      ok = ok_(hipGetLastError()) && DevCopy(g, &f, flags, 4, 2);
   }
   if(ok && f == 0)
   {
      hipLaunchKernelGGL(k_opplan_slabs, dim3((unsigned int)((nchunk + 255) / 256)), dim3(256), 0, s, rlen, nv, nchunk, wcnt, ssum);
      ok = ok_(hipGetLastError()) && ScanCounts(g, hcnt, ntile, *vofs, sum_halo, max_halo)
        && ScanCounts(g, wcnt, nchunk, *noff, nslab, &wmax) && ScanCounts(g, ssum, nchunk, soff, &rs, &wmax);
      *sum_ring = rs;
      if(ok)
      {
         *vlist = (int *)DevAlloc(g, (size_t)*sum_halo * 4 + 4);
         *nbr = (unsigned short *)DevAlloc(g, (size_t)*nslab * 128 + 128);
         ok = *vlist && *nbr && ok_(hipMemsetAsync(*nbr, 0, (size_t)*nslab * 128 + 128, s));
      }
      if(ok)
      {
         hipLaunchKernelGGL(k_opplan_fill, dim3((unsigned int)ntile), dim3(kPB), kOpLds, s, ent, off, deg, (const int4 *)tet,
                            nv, TV, *vofs, *vlist, *noff, *nbr, flags);
         ok = ok_(hipGetLastError()) && ok_(hipStreamSynchronize(s));
      }
      rc = ok ? 1 : -1;
   }
   else if(ok)
      rc = (f & 2) ? -1 : 0;
   DevFree(g, hcnt); DevFree(g, rlen); DevFree(g, wcnt); DevFree(g, ssum); DevFree(g, soff); DevFree(g, flags);
   if(rc != 1)
   {
      DevFree(g, *vofs); DevFree(g, *vlist); DevFree(g, *noff); DevFree(g, *nbr);
      *vofs = nullptr; *vlist = nullptr; *noff = nullptr; *nbr = nullptr;
   }
   return rc;
}

} // namespace gml
