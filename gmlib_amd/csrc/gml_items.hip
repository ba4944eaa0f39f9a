// Topology from element items on the GPU: edge and face extraction
// (GmlExtractEdges / GmlExtractFaces), downlinks (element -> its edges or
// faces), shells (edge or face -> the elements containing it) and
// neighbours (GmlSetNeighbours) of a device handle.
//
// The reference builds all of these with one CPU hash of the items' sorted
// nodes (sources/gmlib3.c:743-1168 links, 3719-4171 extraction, 4178-4267
// neighbours; helpers GetItmNod / CalHshKey / AddHsh / GetHsh,
// 1567-1721).  Here the items (a local edge or face of an element, or an
// entity of an edge / face table) are laid out in the reference's visiting
// order -- source by source, element by element, local item by local item --
// so that an item's index is its visiting rank, and each gets a key: its
// sorted nodes packed b bits each (b from the vertex count) into one 64-bit
// word, or, when deg * b > 64, into a 64-bit low word and a high word.
// A stable LSD radix sort of the item indices by key puts equal items next
// to each other, each run in visiting order (ascending element), and every
// query becomes a run lookup:
//   extraction  the runs' smallest positions, sorted (a block-local pass in
//               LDS first keeps one copy per work-group of each item, with
//               its smallest position, which shrinks the sort 2-4x);
//   downlink    the last item of the run (the highest entity index, the
//               head of the reference's chain);
//   shell       the run reversed (descending element, the chain order);
//   neighbour   the lowest element of the run other than the element itself.

#include "gml_internal.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <climits>
#include <type_traits>
#include <vector>

namespace gml {

namespace {

// the local items of one element kind used by a query, in local order:
// node indices and the local index a link carries for it
struct LocalItems {
   int n;
   int code[12];
   int nod[12][4];
};

struct Source {
   const int *ele;
   long long ne;
   int len;
   LocalItems li;
};

struct KeyForm {
   int deg, b;        // nodes per item, bits per node
   bool two;          // two-word form: the deg * b-bit key split into a low
   int lo_bits, hi_bits;   // 64-bit word and a high word of deg * b - 64 bits
};

int bits_for(long long n)   // bits to hold 0 .. n-1 (at least 1)
{
   int b = 1;
   while(b < 63 && (1ll << b) < n) b++;
   return b;
}

KeyForm MakeKeyForm(int deg, long long nv)
{
   KeyForm kf;
   kf.deg = deg;
   kf.b = bits_for(std::max(2ll, nv));
   kf.two = deg * kf.b > 64;
   kf.lo_bits = kf.two ? 64 : deg * kf.b;
   kf.hi_bits = kf.two ? deg * kf.b - 64 : 0;
   return kf;
}

// items of degree deg of element kind typ (edges: ItmEdgNod, faces:
// ItmFacNod filtered by ItmFacDeg, vertices: the nodes)
LocalItems ItemsOf(int typ, int deg)
{
   LocalItems li = {};
   if(deg == 1)
      for(int j = 0; j < EleNmbNod[typ]; j++, li.n++) { li.code[li.n] = j; li.nod[li.n][0] = j; }
   else if(deg == 2)
      for(int j = 0; j < NmbTpoLnk[typ][GmlEdges]; j++, li.n++)
      {
         li.code[li.n] = j;
         for(int k = 0; k < 2; k++) li.nod[li.n][k] = ItmEdgNod[typ][j][k];
      }
   else
      for(int j = 0; j < 6 && ItmFacDeg[typ][j]; j++)
         if(ItmFacDeg[typ][j] == deg)
         {
            li.code[li.n] = j;
            for(int k = 0; k < 4; k++) li.nod[li.n][k] = ItmFacNod[typ][j][k];
            li.n++;
         }
   return li;
}

// the entities of a table as one item each (an edge or face table)
LocalItems Itself(int deg)
{
   LocalItems li = {};
   li.n = 1;
   for(int k = 0; k < deg; k++) li.nod[0][k] = k;
   return li;
}

int NodesOf(int typ)   // nodes of an item of kind typ
{
   return typ == GmlVertices ? 1 : typ == GmlEdges ? 2 : typ == GmlTriangles ? 3 : 4;
}

__device__ __forceinline__ void sort_nodes(int *a, int k)
{
   for(int i = 1; i < k; i++)
      for(int j = i; j > 0 && a[j - 1] > a[j]; j--) { const int t = a[j]; a[j] = a[j - 1]; a[j - 1] = t; }
}

// key (lo, hi) of item j of element e; false for a node outside [0, nv)
__device__ __forceinline__ bool item_key(const int *__restrict__ ele, long long e, int len, const LocalItems &li,
                                         int j, const KeyForm &kf, long long nv, unsigned long long *lo,
                                         unsigned long long *hi)
{
   int a[4] = {0, 0, 0, 0};
   bool out = false;
   for(int k = 0; k < kf.deg; k++)
   {
      a[k] = ele[e * len + li.nod[j][k]];
      out |= a[k] < 0 || a[k] >= nv;
   }
   sort_nodes(a, kf.deg);
   // the sorted nodes packed b bits each, most significant first, as a
   // (hi, lo) pair of 64-bit words (b < 64)
   unsigned long long l = 0, h = 0;
   for(int i = 0; i < kf.deg; i++)
   {
      h = (h << kf.b) | (l >> (64 - kf.b));
      l = (l << kf.b) | (unsigned)a[i];
   }
   *lo = l;
   *hi = kf.two ? h : 0;
   return !out;
}

// key of item j of element e (index base + e * li.n + j); a node outside
// [0, nv) raises *bad
__global__ void k_item_keys(long long ne, int len, const int *__restrict__ ele, LocalItems li, KeyForm kf,
                            long long nv, unsigned long long *__restrict__ klo, unsigned long long *__restrict__ khi,
                            int *__restrict__ bad)
{
   const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(t >= ne * li.n) return;
   const long long e = t / li.n;
   const int j = (int)(t - e * li.n);
   unsigned long long lo, hi;
   if(!item_key(ele, e, len, li, j, kf, nv, &lo, &hi)) *bad = 1;
   klo[t] = lo;
   if(kf.two) khi[t] = hi;
}

__global__ void k_iota(long long n, int *__restrict__ p)
{
   const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(i < n) p[i] = (int)i;
}

__global__ void k_gather_u64(long long n, const int *__restrict__ ord, const unsigned long long *__restrict__ src,
                             unsigned long long *__restrict__ dst)
{
   const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(i < n) dst[i] = src[ord[i]];
}

__device__ __forceinline__ bool key_less(unsigned long long ahi, unsigned long long alo, unsigned long long bhi,
                                         unsigned long long blo)
{
   return ahi < bhi || (ahi == bhi && alo < blo);
}

// first sorted position whose key is greater than (hi, lo) (upper = true)
// or not less than it (upper = false)
__device__ long long bound(long long n, const unsigned long long *__restrict__ slo,
                           const unsigned long long *__restrict__ shi, unsigned long long hi, unsigned long long lo,
                           bool upper)
{
   long long a = 0, b = n;
   while(a < b)
   {
      const long long m = (a + b) >> 1;
      const unsigned long long mh = shi ? shi[m] : 0, ml = slo[m];
      const bool go_right = upper ? !key_less(hi, lo, mh, ml) : key_less(mh, ml, hi, lo);
      if(go_right) a = m + 1;
      else b = m;
   }
   return a;
}

// Items sorted by key: klo/khi by item index, slo/shi in sorted order, ord
// the item index at each sorted position (khi, shi null in the one-word form)
struct ItemSort {
   GmlCtx *g;
   long long n = 0;
   KeyForm kf;
   unsigned long long *klo = nullptr, *khi = nullptr, *slo = nullptr, *shi = nullptr;
   int *ord = nullptr, *idx = nullptr;
   void *tmp = nullptr;
   size_t tbytes = 0;
   explicit ItemSort(GmlCtx *c) : g(c) {}
   ~ItemSort()
   {
      for(void *p : {(void *)klo, (void *)khi, (void *)slo, (void *)shi, (void *)ord, (void *)idx, tmp}) ItemFree(g, p);
   }
};

template <typename T> bool DAlloc(GmlCtx *g, T **p, size_t bytes) { return ItemAlloc(g, (void **)p, bytes); }

const unsigned kB = 256;
dim3 blocks(long long n) { return dim3((unsigned)((n + kB - 1) / kB)); }
bool ok_(hipError_t e) { return e == hipSuccess; }

} // namespace

// Temporaries of the item kernels (keys, sort buffers: several GB for a
// 100M-tet mesh) are cached per handle: a freed block is kept and handed
// out again to a later request it fits (best fit), so repeated extractions
// and link builds map their memory once instead of at every call (mapping
// and unmapping multi-GB buffers costs more than the kernels).  All users
// run on the handle's stream, so a block freed by one kernel sequence and
// reused by the next is ordered by the stream.  GmlStop releases the cache;
// a failed allocation first releases the idle blocks and retries.
bool ItemAlloc(GmlCtx *g, void **p, size_t bytes)
{
   bytes = (std::max<size_t>(bytes, 16) + 255) & ~(size_t)255;
   ScratchBlock *best = nullptr;
   for(ScratchBlock &b : g->iscratch)
      if(!b.busy && b.bytes >= bytes && (!best || b.bytes < best->bytes)) best = &b;
   if(best)
   {
      best->busy = true;
      *p = best->p;
      return true;
   }
   void *q = nullptr;
   if(hipMalloc(&q, bytes) != hipSuccess)
   {
      (void)hipGetLastError();
      FreeItemPool(g, true);
      if(hipMalloc(&q, bytes) != hipSuccess) { (void)hipGetLastError(); *p = nullptr; return false; }
   }
   g->iscratch.push_back({q, bytes, true});
   *p = q;
   return true;
}

void ItemFree(GmlCtx *g, void *p)
{
   if(!p) return;
   for(ScratchBlock &b : g->iscratch)
      if(b.p == p) { b.busy = false; return; }
}

// release the cached blocks (idle_only: those not handed out)
void FreeItemPool(GmlCtx *g, bool idle_only)
{
   if(g->iscratch.empty()) return;
   (void)hipStreamSynchronize(g->stream);
   std::vector<ScratchBlock> keep;
   for(ScratchBlock &b : g->iscratch)
      if(idle_only && b.busy) keep.push_back(b);
      else (void)hipFree(b.p);
   g->iscratch.swap(keep);
}

namespace {

// the sort buffers of an item sort whose key arrays exist (slo, shi, ord,
// idx and the temporary storage)
bool AllocSortBuffers(GmlCtx *g, ItemSort *it)
{
   hipStream_t s = g->stream;
   const long long n = it->n;
   const KeyForm &kf = it->kf;
   if(!n) return true;
   size_t t1 = 0, t2 = 0, t3 = 0, t4 = 0;
   int *cnt = nullptr;
   unsigned char *flags = nullptr;
   bool ok = DAlloc(g, &it->slo, n * 8) && (!kf.two || DAlloc(g, &it->shi, n * 8))
          && DAlloc(g, &it->ord, n * 4) && DAlloc(g, &it->idx, n * 4)
          && ok_(hipcub::DeviceRadixSort::SortPairs(nullptr, t1, it->klo, it->slo, it->idx, it->ord, (int)n, 0, 64, s))
          && ok_(hipcub::DeviceSelect::Flagged(nullptr, t2, it->ord, flags, it->idx, cnt, (int64_t)n, s))
          && ok_(hipcub::DeviceRadixSort::SortKeys(nullptr, t3, it->ord, it->idx, (int)n, 0, 32, s))
          && ok_(hipcub::DeviceRadixSort::SortPairs(nullptr, t4, it->ord, it->idx, it->ord, it->idx, (int)n, 0, 32, s));
   if(ok)
   {
      it->tbytes = std::max(std::max(t1, t2), std::max(t3, t4)) + 256;
      ok = DAlloc(g, &it->tmp, it->tbytes);
   }
   return ok;
}

// the key arrays of n items (klo, and khi in the two-word form)
bool AllocKeys(GmlCtx *g, long long n, const KeyForm &kf, ItemSort *it)
{
   it->n = n;
   it->kf = kf;
   return !n || (DAlloc(g, &it->klo, n * 8) && (!kf.two || DAlloc(g, &it->khi, n * 8)));
}

// Arrays of an item sort of n items; the temporary storage also fits a
// select, a 32-bit key sort and a 32-bit pair sort over n items.
bool AllocItems(GmlCtx *g, long long n, const KeyForm &kf, ItemSort *it)
{
   return AllocKeys(g, n, kf, it) && AllocSortBuffers(g, it);
}

// stable sort of the it->n keys already in klo / khi
bool SortPrepared(hipStream_t s, ItemSort *it)
{
   const long long n = it->n;
   const KeyForm &kf = it->kf;
   if(!n) return true;
   bool ok = true;
   {
      hipLaunchKernelGGL(k_iota, blocks(n), dim3(kB), 0, s, n, it->idx);
      size_t t = it->tbytes;
      ok = ok_(hipcub::DeviceRadixSort::SortPairs(it->tmp, t, it->klo, it->slo, it->idx, it->ord, (int)n, 0,
                                                  kf.lo_bits, s));
   }
   if(ok && kf.two)
   {
      // stable pass on the high words, carrying the order; then the low
      // words in the final order
      hipLaunchKernelGGL(k_gather_u64, blocks(n), dim3(kB), 0, s, n, it->ord, it->khi, it->slo);
      size_t t = it->tbytes;
      ok = ok_(hipcub::DeviceRadixSort::SortPairs(it->tmp, t, it->slo, it->shi, it->ord, it->idx, (int)n, 0,
                                                  kf.hi_bits, s));
      std::swap(it->ord, it->idx);
      hipLaunchKernelGGL(k_gather_u64, blocks(n), dim3(kB), 0, s, n, it->ord, it->klo, it->slo);
   }
   return ok && ok_(hipGetLastError());
}

// Keys of every item of the sources (in order) and their stable sort.
// bad: device flag raised by a node out of range.
bool SortItems(GmlCtx *g, const std::vector<Source> &src, const KeyForm &kf, long long nv, int *bad,
               ItemSort *it)
{
   hipStream_t s = g->stream;
   long long n = 0;
   for(const Source &sr : src) n += sr.ne * sr.li.n;
   if(!AllocItems(g, n, kf, it)) return false;
   long long at = 0;
   for(const Source &sr : src)
   {
      const long long m = sr.ne * sr.li.n;
      if(m)
         hipLaunchKernelGGL(k_item_keys, blocks(m), dim3(kB), 0, s, sr.ne, sr.len, sr.ele, sr.li, kf, nv,
                            it->klo + at, kf.two ? it->khi + at : nullptr, bad);
      at += m;
   }
   return SortPrepared(s, it);
}

// queries: keys of the items of the sources, unsorted
struct QueryKeys {
   GmlCtx *g;
   long long n = 0;
   unsigned long long *lo = nullptr, *hi = nullptr;
   explicit QueryKeys(GmlCtx *c) : g(c) {}
   ~QueryKeys()
   {
      ItemFree(g, lo);
      ItemFree(g, hi);
   }
};

bool KeysOf(GmlCtx *g, const Source &sr, const KeyForm &kf, long long nv, int *bad, QueryKeys *q)
{
   hipStream_t s = g->stream;
   q->n = sr.ne * sr.li.n;
   if(!q->n) return true;
   if(!DAlloc(g, &q->lo, q->n * 8) || (kf.two && !DAlloc(g, &q->hi, q->n * 8))) return false;
   hipLaunchKernelGGL(k_item_keys, blocks(q->n), dim3(kB), 0, s, sr.ne, sr.len, sr.ele, sr.li, kf, nv, q->lo, q->hi,
                      bad);
   return ok_(hipGetLastError());
}

// ---------------------------------------------------------------------------
// extraction
// ---------------------------------------------------------------------------

// Block-local first appearances: a work-group takes kDedupThreads elements
// of one source (one per thread), inserts their items' keys into an LDS
// hash table that keeps each key's smallest position, and appends its
// distinct items (key, position) to the output, reserved with one global
// atomic.  Items of an element are inserted one local index at a time, in
// rounds separated by barriers: an item claims an empty slot (compare-and-
// swap on the slot's owner word) or, after the barrier, compares its key
// with the slot's (written before the barrier by the slot's owner) and
// either keeps the smaller position there or probes the next slot.  At most
// HS x 3/4 items (HS = 2048 for up to 6 items per element, 4096 for up to
// 12) go into HS slots, so a slot is always free and every round ends.
constexpr int kDedupThreads = 256;

// LDS hash slot of a key: the xor-shift folds the word's high bits down
// before the multiply, whose bits 40.. then depend on every key bit
__device__ __forceinline__ unsigned int dedup_slot(unsigned long long lo, unsigned long long hi)
{
   unsigned long long h = lo ^ (hi * 0xc2b2ae3d27d4eb4full);
   h ^= h >> 31;
   h *= 0x9e3779b97f4a7c15ull;
   return (unsigned int)(h >> 40);
}

template <bool TWO, int HS>
__global__ void __launch_bounds__(kDedupThreads)
k_item_dedup(long long ne, int len, const int *__restrict__ ele, LocalItems li, KeyForm kf, long long nv,
             long long pos0, unsigned long long *__restrict__ olo, unsigned long long *__restrict__ ohi,
             int *__restrict__ opos, unsigned long long *__restrict__ ocnt, int *__restrict__ bad,
             const unsigned char *__restrict__ only)
{
   if(only && !only[blockIdx.x]) return;   // a block the relative-key pass took
   __shared__ unsigned long long hlo[HS];
   __shared__ unsigned long long hhi[TWO ? HS : 1];
   __shared__ int hpos[HS], hown[HS];
   __shared__ unsigned int nuniq, nout;
   __shared__ unsigned long long base;
   for(int i = threadIdx.x; i < HS; i += kDedupThreads) hown[i] = 0;
   if(threadIdx.x == 0) nuniq = nout = 0;
   __syncthreads();
   const long long e = (long long)blockIdx.x * kDedupThreads + threadIdx.x;
   const bool act = e < ne;
   for(int j = 0; j < li.n; j++)
   {
      unsigned long long lo = 0, hi = 0;
      if(act && !item_key(ele, e, len, li, j, kf, nv, &lo, &hi)) *bad = 1;
      const int pos = act ? (int)(pos0 + e * li.n + j) : 0;
      unsigned int sl = dedup_slot(lo, hi) & (HS - 1);
      bool pend = act;
      while(__syncthreads_or(pend))
      {
         bool check = false;
         if(pend)
         {
            if(atomicCAS(&hown[sl], 0, 1) == 0)
            {
               hlo[sl] = lo;
               if(TWO) hhi[sl] = hi;
               hpos[sl] = pos;
               atomicAdd(&nuniq, 1u);
               pend = false;
            }
            else
               check = true;
         }
         __syncthreads();
         if(check)
         {
            if(hlo[sl] == lo && (!TWO || hhi[sl] == hi))
            {
               atomicMin(&hpos[sl], pos);
               pend = false;
            }
            else
               sl = (sl + 1) & (HS - 1);
         }
      }
   }
   __syncthreads();
   if(threadIdx.x == 0) base = atomicAdd(ocnt, (unsigned long long)nuniq);
   __syncthreads();
   for(int i = threadIdx.x; i < HS; i += kDedupThreads)
      if(hown[i])
      {
         const unsigned long long k = base + atomicAdd(&nout, 1u);
         olo[k] = hlo[i];
         if(TWO) ohi[k] = hhi[i];
         opos[k] = hpos[i];
      }
}

// The same block-local pass for one-word keys, without rounds: an item
// claims a slot or finds its key there with one 64-bit compare-and-swap on
// the slot's key (EMPTY marks a free slot: a one-word key of valid nodes
// stays below 2^62), then keeps the smaller position with an atomic min.
// Every lane probes on its own, with no barrier in the loop; the table
// keeps the same load bound, so every probe sequence reaches a free slot or
// its key.  An item with a node out of range raises *bad and is not
// inserted.
constexpr unsigned long long kEmptyKey = ~0ull;

template <int HS>
__global__ void __launch_bounds__(kDedupThreads)
k_item_dedup1(long long ne, int len, const int *__restrict__ ele, LocalItems li, KeyForm kf, long long nv,
              long long pos0, unsigned long long *__restrict__ olo, int *__restrict__ opos,
              unsigned long long *__restrict__ ocnt, int *__restrict__ bad)
{
   __shared__ unsigned long long hkey[HS];
   __shared__ int hpos[HS];
   __shared__ unsigned int nuniq, nout;
   __shared__ unsigned long long base;
   for(int i = threadIdx.x; i < HS; i += kDedupThreads) { hkey[i] = kEmptyKey; hpos[i] = INT_MAX; }
   if(threadIdx.x == 0) nuniq = nout = 0;
   __syncthreads();
   const long long e = (long long)blockIdx.x * kDedupThreads + threadIdx.x;
   if(e < ne)
      for(int j = 0; j < li.n; j++)
      {
         unsigned long long lo = 0, hi = 0;
         if(!item_key(ele, e, len, li, j, kf, nv, &lo, &hi)) { *bad = 1; continue; }
         const int pos = (int)(pos0 + e * li.n + j);
         unsigned int sl = dedup_slot(lo, 0) & (HS - 1);
         for(;;)
         {
            const unsigned long long prev = atomicCAS(&hkey[sl], kEmptyKey, lo);
            if(prev == kEmptyKey) { atomicAdd(&nuniq, 1u); break; }
            if(prev == lo) break;
            sl = (sl + 1) & (HS - 1);
         }
         atomicMin(&hpos[sl], pos);
      }
   __syncthreads();
   if(threadIdx.x == 0) base = atomicAdd(ocnt, (unsigned long long)nuniq);
   __syncthreads();
   for(int i = threadIdx.x; i < HS; i += kDedupThreads)
      if(hkey[i] != kEmptyKey)
      {
         const unsigned long long k = base + atomicAdd(&nout, 1u);
         olo[k] = hkey[i];
         opos[k] = hpos[i];
      }
}

// The compare-and-swap pass for two-word keys: a work-group whose elements'
// nodes span a range r with deg x bits(r) <= 63 packs each item's nodes
// relative to the group's lowest node into one word, deduplicates those as
// k_item_dedup1 does, and writes the items back in the absolute (lo, hi)
// form.  A group whose range is wider (or that holds a node out of range)
// marks itself in defer[] and leaves its elements to the round-based pass,
// launched next with only = defer.  With a Hilbert-ordered numbering almost
// every group fits.
template <int HS>
__global__ void __launch_bounds__(kDedupThreads)
k_item_dedup_rel(long long ne, int len, const int *__restrict__ ele, LocalItems li, KeyForm kf, long long nv,
                 long long pos0, unsigned long long *__restrict__ olo, unsigned long long *__restrict__ ohi,
                 int *__restrict__ opos, unsigned long long *__restrict__ ocnt, unsigned char *__restrict__ defer)
{
   __shared__ unsigned long long hkey[HS];
   __shared__ int hpos[HS];
   __shared__ int nlo, nhi;
   __shared__ unsigned int nuniq, nout;
   __shared__ unsigned long long base;
   for(int i = threadIdx.x; i < HS; i += kDedupThreads) { hkey[i] = kEmptyKey; hpos[i] = INT_MAX; }
   if(threadIdx.x == 0) { nuniq = nout = 0; nlo = INT_MAX; nhi = INT_MIN; }
   __syncthreads();
   const long long e = (long long)blockIdx.x * kDedupThreads + threadIdx.x;
   const bool act = e < ne;
   int mn = INT_MAX, mx = INT_MIN;
   if(act)   // the nodes the items use (a table row may carry padding)
      for(int j = 0; j < li.n; j++)
         for(int k = 0; k < kf.deg; k++)
         {
            const int v = ele[e * len + li.nod[j][k]];
            mn = min(mn, v);
            mx = max(mx, v);
         }
   for(int o = 32; o > 0; o >>= 1)
   {
      mn = min(mn, __shfl_xor(mn, o));
      mx = max(mx, __shfl_xor(mx, o));
   }
   if((threadIdx.x & 63) == 0) { atomicMin(&nlo, mn); atomicMax(&nhi, mx); }
   __syncthreads();
   const int lo0 = nlo, hi0 = nhi;
   if(lo0 > hi0) return;                          // no element in this group
   int rb = 1;
   while(rb < 31 && (1ll << rb) < (long long)hi0 - lo0 + 1) rb++;
   if(lo0 < 0 || hi0 >= nv || kf.deg * rb > 63)
   {
      if(threadIdx.x == 0) defer[blockIdx.x] = 1;
      return;
   }
   if(act)
      for(int j = 0; j < li.n; j++)
      {
         int a[4] = {0, 0, 0, 0};
         for(int k = 0; k < kf.deg; k++) a[k] = ele[e * len + li.nod[j][k]] - lo0;
         sort_nodes(a, kf.deg);
         unsigned long long key = 0;
         for(int k = 0; k < kf.deg; k++) key = (key << rb) | (unsigned)a[k];
         const int pos = (int)(pos0 + e * li.n + j);
         unsigned int sl = dedup_slot(key, 0) & (HS - 1);
         for(;;)
         {
            const unsigned long long prev = atomicCAS(&hkey[sl], kEmptyKey, key);
            if(prev == kEmptyKey) { atomicAdd(&nuniq, 1u); break; }
            if(prev == key) break;
            sl = (sl + 1) & (HS - 1);
         }
         atomicMin(&hpos[sl], pos);
      }
   __syncthreads();
   if(threadIdx.x == 0) base = atomicAdd(ocnt, (unsigned long long)nuniq);
   __syncthreads();
   for(int i = threadIdx.x; i < HS; i += kDedupThreads)
      if(hkey[i] != kEmptyKey)
      {
         // the absolute key: nodes + lo0 packed b bits each, most significant first
         unsigned long long r = hkey[i], l = 0, h = 0;
         int a[4] = {0, 0, 0, 0};
         for(int k = kf.deg - 1; k >= 0; k--) { a[k] = (int)(r & ((1ull << rb) - 1)) + lo0; r >>= rb; }
         for(int k = 0; k < kf.deg; k++)
         {
            h = (h << kf.b) | (l >> (64 - kf.b));
            l = (l << kf.b) | (unsigned)a[k];
         }
         const unsigned long long k = base + atomicAdd(&nout, 1u);
         olo[k] = l;
         ohi[k] = h;
         opos[k] = hpos[i];
      }
}

// for each run of equal keys in sorted order (after the block-local pass a
// key occurs at most once per work-group): its head gets flag 1 and the
// run's smallest position
__global__ void k_run_first(long long n, const unsigned long long *__restrict__ slo,
                            const unsigned long long *__restrict__ shi, const int *__restrict__ ord,
                            const int *__restrict__ pos, unsigned char *__restrict__ head, int *__restrict__ first)
{
   const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(i >= n) return;
   const bool h = i == 0 || slo[i] != slo[i - 1] || (shi && shi[i] != shi[i - 1]);
   head[i] = h;
   if(!h) return;
   int m = pos[ord[i]];
   for(long long k = i + 1; k < n && slo[k] == slo[i] && (!shi || shi[k] == shi[i]); k++) m = min(m, pos[ord[k]]);
   first[i] = m;
}

// row k of the new table: the nodes of item u[k] decoded from its key, its
// ref (the existing table's ref for a first appearance in it, else 0)
__global__ void k_item_write(long long m, const int *__restrict__ u, const int *__restrict__ upos, KeyForm kf,
                             const unsigned long long *__restrict__ klo, const unsigned long long *__restrict__ khi,
                             long long nold, const int *__restrict__ oref, int w, int *__restrict__ tab,
                             int *__restrict__ ref)
{
   const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(k >= m) return;
   const int i = u[k], p = upos[k];
   int a[4] = {0, 0, 0, 0};
   unsigned long long lo = klo[i], hi = kf.two ? khi[i] : 0;
   const unsigned long long mask = (1ull << kf.b) - 1;
   for(int j = kf.deg - 1; j >= 0; j--)
   {
      a[j] = (int)(lo & mask);
      lo = (lo >> kf.b) | (hi << (64 - kf.b));
      hi >>= kf.b;
   }
   for(int j = 0; j < w; j++) tab[k * w + j] = j < kf.deg ? a[j] : 0;
   ref[k] = (p < nold && oref) ? oref[p] : 0;
}

// Bucketed first appearances, the default after the block-local pass: each
// block-distinct item goes to a bucket named by its lowest node, a vertex
// (with a Hilbert-ordered numbering a work-group's buckets lie close
// together), split four ways by its second node, and one counting atomic
// per item gives its slot there; after a scan of the bucket sizes the items
// are copied into bucket order, each flags its position if no item of its
// bucket has an equal key at a smaller position (the item's first
// appearance), and a compaction of all positions lists the flagged ones in
// ascending order: the new table's order.  A bucket of more than kMaxBucket
// items (one vertex under very many items) sends the call down the sorted
// path instead, which has no such bound.
constexpr int kMaxBucket = 256;

// node i of a key (0 = the lowest node: the most significant b bits)
__device__ __forceinline__ int key_node(unsigned long long lo, unsigned long long hi, const KeyForm &kf, int i)
{
   const int sh = kf.b * (kf.deg - 1 - i);
   const unsigned long long v = sh >= 64 ? hi >> (sh - 64) : (lo >> sh) | (sh ? hi << (64 - sh) : 0ull);
   return (int)(v & ((1ull << kf.b) - 1));
}

// the bucket of a key: its lowest node, split kSubBuckets ways by the low
// bits of its second node (equal keys share both)
constexpr int kSubBuckets = 4;
__device__ __forceinline__ long long key_bucket(unsigned long long lo, unsigned long long hi, const KeyForm &kf)
{
   return (long long)key_node(lo, hi, kf, 0) * kSubBuckets + (key_node(lo, hi, kf, 1) & (kSubBuckets - 1));
}

__global__ void k_bucket_count(long long n, const unsigned long long *__restrict__ klo,
                               const unsigned long long *__restrict__ khi, KeyForm kf, int *__restrict__ cnt,
                               int *__restrict__ slot)
{
   const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(i >= n) return;
   slot[i] = atomicAdd(&cnt[key_bucket(klo[i], khi ? khi[i] : 0ull, kf)], 1);
}

// an item in bucket order: its key and position in one record, so that the
// scattered copy into bucket order is one store per item
struct BucketItem1 { unsigned long long lo; int pos, pad; };
struct BucketItem2 { unsigned long long lo, hi; int pos, pad; };
__device__ __forceinline__ unsigned long long item_hi(const BucketItem1 &) { return 0ull; }
__device__ __forceinline__ unsigned long long item_hi(const BucketItem2 &r) { return r.hi; }
__device__ __forceinline__ void set_hi(BucketItem1 &, unsigned long long) {}
__device__ __forceinline__ void set_hi(BucketItem2 &r, unsigned long long hi) { r.hi = hi; }
__device__ __forceinline__ void pad_item(BucketItem1 &r) { r.pad = 0; }
__device__ __forceinline__ void pad_item(BucketItem2 &r) { r.pad = 0; }

template <class R>
__global__ void k_bucket_fill(long long n, const unsigned long long *__restrict__ klo,
                              const unsigned long long *__restrict__ khi, const int *__restrict__ pos,
                              const int *__restrict__ slot, const int *__restrict__ off, KeyForm kf,
                              R *__restrict__ rec)
{
   const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(i >= n) return;
   R r;
   r.lo = klo[i];
   set_hi(r, khi ? khi[i] : 0ull);
   r.pos = pos[i];
   pad_item(r);
   rec[(long long)off[key_bucket(r.lo, item_hi(r), kf)] + slot[i]] = r;
}

// one thread per item in bucket order: flag[p] = 1 for the item at position
// p if no item of its bucket has an equal key at a smaller position.  The
// lanes of a wave scan the same or neighbouring buckets, so the loop's loads
// are shared through the cache.
template <class R>
__global__ void k_bucket_first(long long n, KeyForm kf, const int *__restrict__ cnt, const int *__restrict__ off,
                               const R *__restrict__ rec, unsigned char *__restrict__ flag)
{
   const long long d = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(d >= n) return;
   const R r = rec[d];
   const long long a = key_bucket(r.lo, item_hi(r), kf), b = off[a], e = b + cnt[a];
   for(long long k = b; k < e; k++)
   {
      const R q = rec[k];
      if(q.lo == r.lo && item_hi(q) == item_hi(r) && q.pos < r.pos) return;
   }
   flag[r.pos] = 1;
}

// Stream compaction of the flagged positions, in order: kFlagChunk
// positions per work-group, kFlagChunk / 256 = 16 flag bytes per thread read
// as one 16-byte load.  A count pass, a scan of the per-chunk counts, then
// an emit pass with a work-group scan.
constexpr int kFlagChunk = 4096;

__device__ __forceinline__ int flag_bits(const uint4 &v)
{
   return __popc(v.x) + __popc(v.y) + __popc(v.z) + __popc(v.w);   // each flag byte is 0 or 1
}

__global__ void __launch_bounds__(256) k_flag_count(const uint4 *__restrict__ flag, int *__restrict__ ccnt)
{
   typedef hipcub::BlockReduce<int, 256> Red;
   __shared__ typename Red::TempStorage tmp;
   const int c = flag_bits(flag[(size_t)blockIdx.x * 256 + threadIdx.x]);
   const int t = Red(tmp).Sum(c);
   if(threadIdx.x == 0) ccnt[blockIdx.x] = t;
}

__global__ void __launch_bounds__(256) k_flag_emit(const uint4 *__restrict__ flag, const int *__restrict__ coff,
                                                   int *__restrict__ sel)
{
   typedef hipcub::BlockScan<int, 256> Scan;
   __shared__ typename Scan::TempStorage tmp;
   const uint4 v = flag[(size_t)blockIdx.x * 256 + threadIdx.x];
   int at;
   Scan(tmp).ExclusiveSum(flag_bits(v), at);
   at += coff[blockIdx.x];
   const int p0 = blockIdx.x * kFlagChunk + threadIdx.x * 16;
   const unsigned w[4] = {v.x, v.y, v.z, v.w};
   for(int i = 0; i < 16; i++)
      if((w[i >> 2] >> (8 * (i & 3))) & 0xff) sel[at++] = p0 + i;
}

// First appearances through the vertex -> tet ball, for a mesh whose only
// element kind is the tet and that has no table of the extracted kind yet:
// item j of tet e (an edge or a face) appears first at position e * n + j
// iff no tet of smaller index holds all its nodes.  Every tet holding the
// item holds its lowest node a, so the ball of a holds them all.  The thread
// of vertex a walks its ball once in ascending tet index (the ball is stored
// in descending order) and keeps the items whose lowest node is a that it has
// met, as keys in registers (the other node of an edge; the two other nodes
// of a face, lower one in the high word): an item is first iff its key is
// new (the list is searched 8 keys at a time, up to the keys in use).  The
// low four bits of a ball entry are a's local index la in the tet, so the
// candidates of a tet are found without a search: the edges {a, y} and the
// faces opposite each other node y, item numbers from map (edges: nibble
// la * 4 + o for locals la, o; faces: nibble o for the face without local o).
// Output: byte e * 4 + la of mark (zeroed by the caller) gets bit j for each
// first item j of tet e whose lowest node is a, written only when some bit
// is set (at most one store per first item, where a byte per item and
// position costs a scattered store per item); every item has one lowest
// node, so the OR of a tet's four bytes is the set of its first items.  A
// vertex with more than S distinct items splits them into P classes (key
// hash mod P) and walks its ball once per class, P doubling until each class
// fits.  A tet that repeats a vertex, a ball entry that does not name a, or
// a class still too large at kBallMaxPass classes sets *fallback (the
// general path then runs).
constexpr int kBallMaxDeg = 256;   // wider balls take the general path
constexpr int kBallMaxPass = 64;

template <typename T, int S>
__device__ __forceinline__ bool seen_has(const T (&seen)[S], int cnt, T key)
{
   bool found = false;
#pragma unroll
   for(int c0 = 0; c0 < S; c0 += 8)
      if(c0 < cnt)
#pragma unroll
         for(int s = c0; s < c0 + 8; s++) found |= seen[s] == key;
   return found;
}

// insert at the front (cnt < S entries in use; entries from cnt on hold ~0,
// which no key equals): every index is a constant, so the list stays in
// registers; the chunks above the ones in use are not moved
template <typename T, int S>
__device__ __forceinline__ void seen_put(T (&seen)[S], int cnt, T key)
{
#pragma unroll
   for(int c0 = S - 8; c0 >= 0; c0 -= 8)
      if(c0 <= cnt)
#pragma unroll
         for(int s = c0 + 7; s >= c0; s--) seen[s] = s == 0 ? key : seen[s - 1];
}

template <int KD, int S>
__global__ void __launch_bounds__(256) k_item_first_ball(long long nv, const int *__restrict__ deg,
                                                         const unsigned int *__restrict__ off,
                                                         const int *__restrict__ ent, const int4 *__restrict__ tet,
                                                         unsigned long long map, unsigned char *__restrict__ mark,
                                                         int *fallback)
{
   // work-groups go to the eight XCDs in turn: each XCD takes one contiguous
   // eighth of the vertices (the grid is a multiple of 8), so the mark lines
   // of a range of tets are written in one L2
   const unsigned int per = gridDim.x / 8, vb = (blockIdx.x % 8) * per + blockIdx.x / 8;
   const long long a = (long long)vb * blockDim.x + threadIdx.x;
   if(a >= nv) return;
   const int d = deg[a], ia = (int)a;
   const size_t base = (size_t)off[a >> 6] * 64 + (a & 63);
   for(int P = 1; P <= kBallMaxPass; P <<= 1)
   {
      bool over = false;
      for(int p = 0; p < P && !over; p++)
      {
         typedef typename std::conditional<KD == 2, unsigned int, unsigned long long>::type Key;
         Key seen[S];
#pragma unroll
         for(int s = 0; s < S; s++) seen[s] = ~(Key)0;
         int cnt = 0;
         int wn = d > 0 ? ent[base + (size_t)(d - 1) * 64] : 0;
         int4 qn = d > 0 ? tet[wn >> 4] : make_int4(0, 0, 0, 0);
         for(int k = d - 1; k >= 0 && !over; k--)
         {
            const int w = wn;
            const int4 q = qn;
            if(k > 0)
            {
               wn = ent[base + (size_t)(k - 1) * 64];
               qn = tet[wn >> 4];
            }
            const long long e = w >> 4;
            const int la = w & 15;
            const int xa = la == 0 ? q.x : la == 1 ? q.y : la == 2 ? q.z : q.w;
            if(la > 3 || xa != ia || q.x == q.y || q.x == q.z || q.x == q.w || q.y == q.z || q.y == q.w || q.z == q.w)
            {
               atomicOr(fallback, 1);
               return;
            }
            // the other three nodes in local order, and their local indices
            const int y[3] = {la == 0 ? q.y : q.x, la <= 1 ? q.z : q.y, la <= 2 ? q.w : q.z};
            const int o[3] = {la == 0 ? 1 : 0, la <= 1 ? 2 : 1, la <= 2 ? 3 : 2};
            unsigned int bits = 0;
#pragma unroll
            for(int i = 0; i < 3; i++)
            {
               Key key;
               int j;
               bool mine;
               if constexpr(KD == 2)
               {
                  key = (unsigned int)y[i];
                  j = (int)((map >> (4 * (la * 4 + o[i]))) & 15);
                  mine = y[i] > ia && ((unsigned int)y[i] & (unsigned int)(P - 1)) == (unsigned int)p;
               }
               else
               {
                  const int b = y[(i + 1) % 3], c = y[(i + 2) % 3];
                  const int lo = min(b, c), hi = max(b, c);
                  key = ((unsigned long long)(unsigned int)lo << 32) | (unsigned int)hi;
                  j = (int)((map >> (4 * o[i])) & 15);
                  mine = lo > ia && (((unsigned int)lo + (unsigned int)hi) & (unsigned int)(P - 1)) == (unsigned int)p;
               }
               if(mine && !over)
               {
                  if(!seen_has<Key, S>(seen, cnt, key))
                  {
                     if(cnt == S)
                        over = true;
                     else
                     {
                        seen_put<Key, S>(seen, cnt++, key);
                        bits |= 1u << j;
                     }
                  }
               }
            }
            // a set bit is a first item whatever happens later in the walk;
            // a walk split into classes adds its bits to the earlier ones
            // (the byte has one writer: the thread of a)
            if(bits)
            {
               unsigned char *mb = &mark[e * 4 + la];
               *mb = (unsigned char)(P == 1 ? bits : *mb | bits);
            }
         }
      }
      if(!over) return;
   }
   atomicOr(fallback, 1);
}

// item numbers of the tet's local edges (deg 2: nibble la * 4 + o for each
// pair of distinct locals) or faces (deg 3: nibble o for the face without
// local o) from its local item table; false when the table is not the six
// edges or four triangles of a tet
static bool BallItemMap(const LocalItems &li, int deg, unsigned long long *map)
{
   unsigned long long m = 0;
   unsigned int set = 0;
   if(deg == 2 && li.n == 6)
      for(int j = 0; j < 6; j++)
      {
         const int p = li.nod[j][0], q = li.nod[j][1];
         if(p < 0 || p > 3 || q < 0 || q > 3 || p == q) return false;
         for(int t : {p * 4 + q, q * 4 + p})
         {
            if(set & (1u << t)) return false;
            set |= 1u << t;
            m |= (unsigned long long)j << (4 * t);
         }
      }
   else if(deg == 3 && li.n == 4)
      for(int j = 0; j < 4; j++)
      {
         unsigned int has = 0;
         for(int i = 0; i < 3; i++)
            if(li.nod[j][i] >= 0 && li.nod[j][i] <= 3) has |= 1u << li.nod[j][i];
         if(__builtin_popcount(has) != 3) return false;
         const int r = __builtin_ctz(~has & 15u);
         if(set & (1u << r)) return false;
         set |= 1u << r;
         m |= (unsigned long long)j << (4 * r);
      }
   else
      return false;
   *map = m;
   return true;
}

// Compaction of the first items marked by k_item_first_ball, in position
// order: kTetChunk tets per work-group, two per thread (one 8-byte load of
// their mark words; the marks past the last tet are zero).  A count pass, a
// scan of the per-chunk counts, then the write pass: each first item's row,
// its nodes sorted ascending (the row k_item_write_pos decodes from the
// key), W columns with zeros past the item's nodes, ref 0.  The write pass
// builds the work-group's rows in LDS and stores them as whole consecutive
// rows.
constexpr int kTetChunk = 512;

__device__ __forceinline__ unsigned int tet_marks(unsigned int w)
{
   return (w | w >> 8 | w >> 16 | w >> 24) & 0xffu;
}

__device__ __forceinline__ int tet_node(const int4 &q, int i)
{
   return i == 0 ? q.x : i == 1 ? q.y : i == 2 ? q.z : q.w;
}

__global__ void __launch_bounds__(256) k_tmark_count(const uint2 *__restrict__ mark, int *__restrict__ ccnt)
{
   typedef hipcub::BlockReduce<int, 256> Red;
   __shared__ typename Red::TempStorage tmp;
   const uint2 v = mark[(size_t)blockIdx.x * 256 + threadIdx.x];
   const int t = Red(tmp).Sum(__popc(tet_marks(v.x)) + __popc(tet_marks(v.y)));
   if(threadIdx.x == 0) ccnt[blockIdx.x] = t;
}

template <int W, int NI, int CAP>
__global__ void __launch_bounds__(256) k_tmark_write(const uint2 *__restrict__ mark, const int *__restrict__ coff,
                                                     const int4 *__restrict__ tet, LocalItems li, int deg,
                                                     int *__restrict__ tab, int *__restrict__ ref)
{
   typedef hipcub::BlockScan<int, 256> Scan;
   typedef typename std::conditional<W == 4, int4, int2>::type Row;
   __shared__ typename Scan::TempStorage tmp;
   __shared__ Row stage[CAP];
   const uint2 v = mark[(size_t)blockIdx.x * 256 + threadIdx.x];
   const unsigned int m[2] = {tet_marks(v.x), tet_marks(v.y)};
   int at, tot;
   Scan(tmp).ExclusiveSum(__popc(m[0]) + __popc(m[1]), at, tot);
   const long long e0 = (long long)blockIdx.x * kTetChunk + threadIdx.x * 2;
   const size_t base = (size_t)coff[blockIdx.x];
   Row *out = reinterpret_cast<Row *>(tab) + base;
   // the stage holds CAP rows (a tet chunk has about two first faces or 1.2
   // first edges per tet); a chunk with more writes its rows straight out
   const bool staged = tot <= CAP;
   for(int t = 0; t < 2; t++)
   {
      if(!m[t]) continue;
      const int4 q = tet[e0 + t];
      for(int j = 0; j < NI; j++)
      {
         if(!((m[t] >> j) & 1u)) continue;
         int a[4] = {0, 0, 0, 0};
         for(int k = 0; k < deg; k++) a[k] = tet_node(q, li.nod[j][k]);
         sort_nodes(a, deg);
         Row r;
         if constexpr(W == 4)
            r = make_int4(a[0], a[1], deg > 2 ? a[2] : 0, deg > 3 ? a[3] : 0);
         else
            r = make_int2(a[0], a[1]);
         if(staged)
            stage[at] = r;
         else
            out[at] = r;
         at++;
      }
   }
   if(staged)
   {
      __syncthreads();
      for(int r = threadIdx.x; r < tot; r += 256) out[r] = stage[r];
   }
   for(int r = threadIdx.x; r < tot; r += 256) ref[base + r] = 0;
}

// the sources of an extraction and the first item position of each
constexpr int kMaxSources = 8;
struct SourceSet {
   int n;
   long long at[kMaxSources];
   Source s[kMaxSources];
};

// row k of the new table from the item at position sel[k]: its element's
// nodes, sorted (decoded from the key as k_item_write does), its ref (the
// existing table's ref for a position in it, else 0)
__global__ void k_item_write_pos(long long m, const int *__restrict__ sel, SourceSet ss, KeyForm kf, long long nv,
                                 long long nold, const int *__restrict__ oref, int w, int *__restrict__ tab,
                                 int *__restrict__ ref)
{
   const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(k >= m) return;
   const long long p = sel[k];
   int q = 0;
   while(q + 1 < ss.n && p >= ss.at[q + 1]) q++;
   const Source &sr = ss.s[q];
   const long long e = (p - ss.at[q]) / sr.li.n;
   const int j = (int)(p - ss.at[q] - e * sr.li.n);
   unsigned long long lo, hi;
   (void)item_key(sr.ele, e, sr.len, sr.li, j, kf, nv, &lo, &hi);
   int a[4] = {0, 0, 0, 0};
   const unsigned long long mask = (1ull << kf.b) - 1;
   for(int i = kf.deg - 1; i >= 0; i--)
   {
      a[i] = (int)(lo & mask);
      lo = (lo >> kf.b) | (hi << (64 - kf.b));
      hi >>= kf.b;
   }
   for(int i = 0; i < w; i++) tab[k * w + i] = i < kf.deg ? a[i] : 0;
   ref[k] = (p < nold && oref) ? oref[p] : 0;
}

// ---------------------------------------------------------------------------
// downlinks, shells, neighbours
// ---------------------------------------------------------------------------

// downlink: row e, column code(j) = the highest entity index with the
// nodes of local item j of element e, 0 if none
__global__ void k_down_lookup(long long nq, int nloc, LocalItems li, const unsigned long long *__restrict__ qlo,
                              const unsigned long long *__restrict__ qhi, long long n,
                              const unsigned long long *__restrict__ slo, const unsigned long long *__restrict__ shi,
                              const int *__restrict__ ord, int rw, int *__restrict__ link)
{
   const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(q >= nq) return;
   const long long e = q / nloc;
   const int j = (int)(q - e * nloc), c = li.code[j];
   if(c >= rw) return;
   const unsigned long long hi = qhi ? qhi[q] : 0, lo = qlo[q];
   const long long ub = bound(n, slo, shi, hi, lo, true);
   int r = 0;
   if(ub > 0 && slo[ub - 1] == lo && (!shi || shi[ub - 1] == hi)) r = ord[ub - 1];
   link[e * rw + c] = r;
}

// shell: the run of each source entity (first sorted position, degree)
__global__ void k_shell_runs(long long ns, const unsigned long long *__restrict__ qlo,
                             const unsigned long long *__restrict__ qhi, long long n,
                             const unsigned long long *__restrict__ slo, const unsigned long long *__restrict__ shi,
                             int *__restrict__ beg, int *__restrict__ deg)
{
   const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(q >= ns) return;
   const unsigned long long hi = qhi ? qhi[q] : 0, lo = qlo[q];
   const long long lb = bound(n, slo, shi, hi, lo, false), ub = bound(n, slo, shi, hi, lo, true);
   beg[q] = (int)lb;
   deg[q] = (int)(ub - lb);
}

// shell rows in SELL-64: slot k of row q = the k-th item of its run from the
// end, as (element << 4) | local index
__global__ void k_shell_fill(long long ns, const int *__restrict__ beg, const int *__restrict__ deg,
                             const int *__restrict__ ord, int nloc, LocalItems li, const unsigned int *__restrict__ off,
                             int *__restrict__ ent)
{
   const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(q >= ns) return;
   const int d = deg[q], b = beg[q];
   int *row = ent + (size_t)off[q >> 6] * 64 + (q & 63);
   for(int k = 0; k < d; k++)
   {
      const int i = ord[b + d - 1 - k], e = i / nloc, j = i - e * nloc;
      row[(size_t)k * 64] = (e << 4) | li.code[j];
   }
}

// neighbours: for sorted position p (item i = ord[p] = element e, local j),
// the lowest element of p's run other than e, 0 if none
__global__ void k_neighbours(long long n, const unsigned long long *__restrict__ slo,
                             const unsigned long long *__restrict__ shi, const int *__restrict__ ord, int nloc,
                             LocalItems li, int rw, int *__restrict__ ngb)
{
   const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(p >= n) return;
   const unsigned long long hi = shi ? shi[p] : 0, lo = slo[p];
   const int i = ord[p], e = i / nloc, j = i - e * nloc;
   // run start: the first position with this key
   const long long rs = bound(p + 1, slo, shi, hi, lo, false);
   int r = 0;
   for(long long k = rs; k < n && slo[k] == lo && (!shi || shi[k] == hi); k++)
   {
      const int f = ord[k] / nloc;
      if(f != e) { r = f; break; }
   }
   if(li.code[j] < rw) ngb[(long long)e * rw + li.code[j]] = r;
}

} // namespace

// Extract the distinct items of degree deg (2: edges, 3: triangles, 4:
// quads) into the table of kind typ, replacing it.  Returns the number of
// items, or -1 on failure (message printed).
int GpuExtractItems(size_t GmlIdx, GmlCtx *g, int typ, int deg)
{
   std::vector<Source> src;
   std::vector<int> kinds;   // element kind of each source (-1: the existing table)
   long long total = 0, nold = 0;
   const int old = g->TypIdx[typ], oref = g->RefIdx[typ];
   if(old)
   {
      const Data *d = &g->dat[old];
      src.push_back({(const int *)d->dev, d->NmbLin, d->ItmLen, Itself(deg)});
      kinds.push_back(-1);
      nold = d->NmbLin;
   }
   const int t0 = deg == 2 ? GmlEdges + 1 : GmlTetrahedra, t1 = deg == 2 ? GmlMaxEleTyp - 1 : GmlHexahedra;
   for(int t = t0; t <= t1; t++)
   {
      if(!g->TypIdx[t] || !g->NmbEle[t]) continue;
      const Data *d = &g->dat[g->TypIdx[t]];
      const Source s = {(const int *)d->dev, d->NmbLin, d->ItmLen, ItemsOf(t, deg)};
      if(s.li.n) { src.push_back(s); kinds.push_back(t); }
   }
   for(const Source &s : src) total += s.ne * s.li.n;
   if(total >= INT_MAX)
   {
      fprintf(stderr, "gml: extraction of %lld items exceeds the 2^31 item limit\n", total);
      return -1;
   }
   if(!total)
   {
      // as the host path: an existing empty table is dropped, none created
      if(old) { GmlFreeData(GmlIdx, old); if(oref && g->dat[oref].use) FreeDataSlot(g, oref); g->RefIdx[typ] = 0; }
      return 0;
   }

   hipStream_t s = g->stream;
   const long long nv = g->NmbEle[GmlVertices];
   const KeyForm kf = MakeKeyForm(deg, nv);
   // the sources with the position of their first item (bucketed path)
   SourceSet ss = {};
   ss.n = (int)std::min(src.size(), (size_t)kMaxSources);   // at most 7: the table and six element kinds
   long long sat = 0;
   for(int i = 0; i < ss.n; i++)
   {
      ss.at[i] = sat;
      ss.s[i] = src[i];
      sat += src[i].ne * src[i].li.n;
   }
   const long long nbk = nv * kSubBuckets;   // buckets: an item's lowest node, split by its second
   auto clock = []() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); };
   double tim[5] = {clock(), 0, 0, 0, 0};
   ItemSort it(g);
   unsigned long long *cnt = nullptr, hc[2] = {0, 0};
   int *pos = nullptr, *first = nullptr, *hsel = nullptr, *usel = nullptr, *p2 = nullptr, *u2 = nullptr, *nsel = nullptr;
   int *bcnt = nullptr, *boff = nullptr, *slot = nullptr, *sel = nullptr, *ccnt = nullptr, *coff = nullptr;
   unsigned char *head = nullptr, *flag = nullptr, *defer = nullptr;
   void *xtmp = nullptr, *brec = nullptr;
   int rc = -1, hbad = 0, m = 0;
   const long long nchunk = (total + kFlagChunk - 1) / kFlagChunk;
   // 0. a mesh of tets only and no table of the kind yet: first appearances
   // through the vertex -> tet ball (built if needed), marked per tet by
   // k_item_first_ball, counted, and the new table written from the marks;
   // the general path below otherwise, and when the ball cannot be built, is
   // wider than kBallMaxDeg or a tet repeats a vertex.  A ball kept from an
   // earlier call is rebuilt when the tet table or the vertex count has
   // changed since; one that a compiled kernel holds is left as it is and
   // the general path runs
   bool viaball = false;
   int ballidx = 0;
   unsigned long long bmap = 0;
   if(!old && src.size() == 1 && kinds[0] == GmlTetrahedra && src[0].len == 4 && BallItemMap(src[0].li, deg, &bmap)
   && !g->host_only && CurrentUplink(g, GmlVertices, GmlTetrahedra)
   && g->up[GmlVertices][GmlTetrahedra].MaxDeg <= kBallMaxDeg)
   {
      const Uplink &u = g->up[GmlVertices][GmlTetrahedra];
      const long long ntc = (src[0].ne + kTetChunk - 1) / kTetChunk;
      const size_t mbytes = (size_t)ntc * kTetChunk * 4;
      int *fb = nullptr, hfb = 1;
      unsigned char *mark = nullptr;
      size_t t3 = 0;
      bool fok = ntc > 0 && DAlloc(g, &fb, 4) && ok_(hipMemsetAsync(fb, 0, 4, s)) && DAlloc(g, &mark, mbytes)
              && ok_(hipMemsetAsync(mark, 0, mbytes, s))
              && DAlloc(g, &ccnt, (ntc + 1) * 4) && DAlloc(g, &coff, (ntc + 1) * 4)
              && ok_(hipMemsetAsync(ccnt + ntc, 0, 4, s))
              && ok_(hipcub::DeviceScan::ExclusiveSum(nullptr, t3, ccnt, coff, (int)(ntc + 1), s))
              && DAlloc(g, &xtmp, t3);
      tim[1] = clock();
      if(fok)
      {
         const dim3 grid((unsigned int)((nv + 255) / 256 + 7) / 8 * 8);
         const int *bdeg = (const int *)g->dat[u.DegIdx].dev, *bent = (const int *)g->dat[u.BalIdx].dev;
         const unsigned int *boff = (const unsigned int *)g->dat[u.OffIdx].dev;
         if(deg == 2)
            hipLaunchKernelGGL((k_item_first_ball<2, 16>), grid, dim3(256), 0, s, nv, bdeg, boff, bent,
                               (const int4 *)src[0].ele, bmap, mark, fb);
         else
            hipLaunchKernelGGL((k_item_first_ball<3, 32>), grid, dim3(256), 0, s, nv, bdeg, boff, bent,
                               (const int4 *)src[0].ele, bmap, mark, fb);
         fok = ok_(hipGetLastError()) && ok_(hipMemcpyAsync(&hfb, fb, 4, hipMemcpyDeviceToHost, s))
            && ok_(hipStreamSynchronize(s));
      }
      tim[2] = clock();
      if(fok && !hfb)
      {
         hipLaunchKernelGGL(k_tmark_count, dim3((unsigned)ntc), dim3(256), 0, s, (const uint2 *)mark, ccnt);
         size_t t = t3;
         fok = ok_(hipGetLastError()) && ok_(hipcub::DeviceScan::ExclusiveSum(xtmp, t, ccnt, coff, (int)(ntc + 1), s))
            && ok_(hipMemcpyAsync(&m, coff + ntc, 4, hipMemcpyDeviceToHost, s)) && ok_(hipStreamSynchronize(s));
         ballidx = fok ? NewMeshTable(g, typ, m, false) : 0;
         if(ballidx)
         {
            const Data *d = &g->dat[ballidx];
            int *tab = (int *)d->dev, *ref = (int *)g->dat[g->RefIdx[typ]].dev;
            if(deg == 2 && d->ItmLen == 2)
               hipLaunchKernelGGL((k_tmark_write<2, 6, 1024>), dim3((unsigned)ntc), dim3(256), 0, s, (const uint2 *)mark,
                                  coff, (const int4 *)src[0].ele, src[0].li, deg, tab, ref);
            else if(deg == 3 && d->ItmLen == 4)
               hipLaunchKernelGGL((k_tmark_write<4, 4, 1536>), dim3((unsigned)ntc), dim3(256), 0, s, (const uint2 *)mark,
                                  coff, (const int4 *)src[0].ele, src[0].li, deg, tab, ref);
            else
               fok = false;
            fok = fok && ok_(hipGetLastError()) && ok_(hipStreamSynchronize(s));
            if(!fok)
            {
               GmlFreeData(GmlIdx, ballidx);
               ballidx = 0;
            }
         }
         viaball = ballidx != 0;
      }
      ItemFree(g, fb);
      ItemFree(g, mark);
      if(!viaball)
      {
         (void)hipGetLastError();
         for(void *q : {(void *)ccnt, (void *)coff, xtmp}) ItemFree(g, q);
         ccnt = nullptr; coff = nullptr; xtmp = nullptr;
         m = 0;
      }
   }
   bool bucketed = false;
   bool ok = true;
   long long n2 = 0;
   if(!viaball)
   {
   // 1. block-local first appearances of every source's items
   long long maxgrid = 1;
   for(const Source &sr : src) maxgrid = std::max(maxgrid, (sr.ne + kDedupThreads - 1) / kDedupThreads);
   ok = DAlloc(g, &cnt, 16) && ok_(hipMemsetAsync(cnt, 0, 16, s)) && DAlloc(g, &nsel, 4)
          && AllocKeys(g, total, kf, &it) && DAlloc(g, &pos, total * 4) && DAlloc(g, &defer, maxgrid);
   tim[1] = clock();
   long long at = 0;
   for(const Source &sr : src)
   {
      if(!ok) break;
      const dim3 grid((unsigned)((sr.ne + kDedupThreads - 1) / kDedupThreads));
      if(!sr.ne) continue;
      const bool small = sr.li.n <= 6;
      if(kf.two)
      {
         // relative one-word keys where a group's node range allows, rounds
         // for the groups that marked themselves
         ok = ok_(hipMemsetAsync(defer, 0, grid.x, s));
         if(!ok) break;
         hipLaunchKernelGGL((small ? k_item_dedup_rel<2048> : k_item_dedup_rel<4096>), grid, dim3(kDedupThreads), 0,
                            s, sr.ne, sr.len, sr.ele, sr.li, kf, nv, at, it.klo, it.khi, pos, cnt, defer);
         hipLaunchKernelGGL((small ? k_item_dedup<true, 2048> : k_item_dedup<true, 4096>), grid, dim3(kDedupThreads),
                            0, s, sr.ne, sr.len, sr.ele, sr.li, kf, nv, at, it.klo, it.khi, pos, cnt, (int *)(cnt + 1),
                            (const unsigned char *)defer);
      }
      else if(kf.deg * kf.b >= 64)   // a full 64-bit key: every value may be a key
         hipLaunchKernelGGL((small ? k_item_dedup<false, 2048> : k_item_dedup<false, 4096>), grid,
                            dim3(kDedupThreads), 0, s, sr.ne, sr.len, sr.ele, sr.li, kf, nv, at, it.klo, nullptr, pos,
                            cnt, (int *)(cnt + 1), nullptr);
      else
         hipLaunchKernelGGL((small ? k_item_dedup1<2048> : k_item_dedup1<4096>), grid, dim3(kDedupThreads), 0, s,
                            sr.ne, sr.len, sr.ele, sr.li, kf, nv, at, it.klo, pos, cnt, (int *)(cnt + 1));
      at += sr.ne * sr.li.n;
   }
   ok = ok && ok_(hipGetLastError()) && ok_(hipMemcpyAsync(hc, cnt, 16, hipMemcpyDeviceToHost, s))
     && ok_(hipStreamSynchronize(s));
   hbad = (int)(hc[1] & 0xffffffffu);
   tim[2] = clock();
   if(ok && hbad)
   {
      fprintf(stderr, "gml: extraction: an element references a vertex outside 0..%lld\n", nv - 1);
      ok = false;
   }
   n2 = (long long)hc[0];
   // 2. bucketed path: the block-distinct items by lowest node, each bucket
   // scanned for its keys' smallest positions, those positions selected in
   // order; the table rows are rebuilt from the elements at them
   size_t tb = 0;
   if(ok && nbk < INT_MAX)
   {
      // one temporary for the bucket reduce and the two scans
      size_t t1 = 0, t2 = 0, t3 = 0;
      int hmax = 0;
      ok = DAlloc(g, &bcnt, nbk * 4) && DAlloc(g, &boff, nbk * 4) && DAlloc(g, &slot, n2 * 4)
        && DAlloc(g, &flag, nchunk * kFlagChunk) && DAlloc(g, &sel, n2 * 4) && DAlloc(g, &ccnt, (nchunk + 1) * 4)
        && DAlloc(g, &coff, (nchunk + 1) * 4)
        && DAlloc(g, &brec, n2 * (kf.two ? sizeof(BucketItem2) : sizeof(BucketItem1)))
        && ok_(hipcub::DeviceReduce::Max(nullptr, t1, bcnt, nsel, (int)nbk, s))
        && ok_(hipcub::DeviceScan::ExclusiveSum(nullptr, t2, bcnt, boff, (int)nbk, s))
        && ok_(hipcub::DeviceScan::ExclusiveSum(nullptr, t3, ccnt, coff, (int)(nchunk + 1), s));
      tb = std::max(t1, std::max(t2, t3));
      ok = ok && DAlloc(g, &xtmp, tb) && ok_(hipMemsetAsync(bcnt, 0, nbk * 4, s));
      if(ok)
      {
         hipLaunchKernelGGL(k_bucket_count, blocks(n2), dim3(kB), 0, s, n2, it.klo, kf.two ? it.khi : nullptr, kf,
                            bcnt, slot);
         size_t t = tb;
         ok = ok_(hipGetLastError()) && ok_(hipcub::DeviceReduce::Max(xtmp, t, bcnt, nsel, (int)nbk, s))
           && ok_(hipMemcpyAsync(&hmax, nsel, 4, hipMemcpyDeviceToHost, s)) && ok_(hipStreamSynchronize(s));
      }
      bucketed = ok && hmax <= kMaxBucket;
   }
   if(bucketed)
   {
      size_t t = tb;
      ok = ok_(hipcub::DeviceScan::ExclusiveSum(xtmp, t, bcnt, boff, (int)nbk, s))
        && ok_(hipMemsetAsync(flag, 0, nchunk * kFlagChunk, s)) && ok_(hipMemsetAsync(ccnt + nchunk, 0, 4, s));
      if(ok)
      {
         if(kf.two)
         {
            hipLaunchKernelGGL(k_bucket_fill<BucketItem2>, blocks(n2), dim3(kB), 0, s, n2, it.klo, it.khi, pos, slot,
                               boff, kf, (BucketItem2 *)brec);
            hipLaunchKernelGGL(k_bucket_first<BucketItem2>, blocks(n2), dim3(kB), 0, s, n2, kf, bcnt, boff,
                               (const BucketItem2 *)brec, flag);
         }
         else
         {
            hipLaunchKernelGGL(k_bucket_fill<BucketItem1>, blocks(n2), dim3(kB), 0, s, n2, it.klo, nullptr, pos, slot,
                               boff, kf, (BucketItem1 *)brec);
            hipLaunchKernelGGL(k_bucket_first<BucketItem1>, blocks(n2), dim3(kB), 0, s, n2, kf, bcnt, boff,
                               (const BucketItem1 *)brec, flag);
         }
         hipLaunchKernelGGL(k_flag_count, dim3((unsigned)nchunk), dim3(256), 0, s, (const uint4 *)flag, ccnt);
         t = tb;
         ok = ok_(hipGetLastError()) && ok_(hipcub::DeviceScan::ExclusiveSum(xtmp, t, ccnt, coff, (int)(nchunk + 1), s));
      }
      if(ok)
      {
         hipLaunchKernelGGL(k_flag_emit, dim3((unsigned)nchunk), dim3(256), 0, s, (const uint4 *)flag, coff, sel);
         ok = ok_(hipGetLastError()) && ok_(hipMemcpyAsync(&m, coff + nchunk, 4, hipMemcpyDeviceToHost, s))
           && ok_(hipStreamSynchronize(s));
      }
   }
   // 2'. sorted path: sort the block-distinct items by key; each run's
   // smallest position is the item's first appearance; those sorted by
   // position are the table
   else if(ok)
   {
      it.n = n2;
      ok = AllocSortBuffers(g, &it) && SortPrepared(s, &it) && DAlloc(g, &head, n2) && DAlloc(g, &first, n2 * 4) && DAlloc(g, &hsel, n2 * 4)
        && DAlloc(g, &usel, n2 * 4) && DAlloc(g, &p2, n2 * 4) && DAlloc(g, &u2, n2 * 4);
      if(ok)
      {
         hipLaunchKernelGGL(k_run_first, blocks(n2), dim3(kB), 0, s, n2, it.slo, it.shi, it.ord, pos, head, first);
         size_t t = it.tbytes;
         ok = ok_(hipGetLastError()) && ok_(hipcub::DeviceSelect::Flagged(it.tmp, t, first, head, hsel, nsel, (int64_t)n2, s));
         t = it.tbytes;
         ok = ok && ok_(hipcub::DeviceSelect::Flagged(it.tmp, t, it.ord, head, usel, nsel, (int64_t)n2, s))
           && ok_(hipMemcpyAsync(&m, nsel, 4, hipMemcpyDeviceToHost, s)) && ok_(hipStreamSynchronize(s));
      }
      if(ok)
      {
         size_t t = it.tbytes;
         ok = ok_(hipcub::DeviceRadixSort::SortPairs(it.tmp, t, hsel, p2, usel, u2, m, 0, bits_for(total), s));
      }
   }
   }
   tim[3] = clock();
   // 3. the new table (created before the old one is dropped: the old refs
   // are read), not zero-filled: the row kernels write every row and ref;
   // the vertex-ball path has written its table already
   if(ok && viaball)
      rc = m;
   else if(ok)
   {
      const int nidx = NewMeshTable(g, typ, m, false);
      if(!nidx) ok = false;
      else
      {
         const Data *d = &g->dat[nidx];
         const int *orf = oref ? (const int *)g->dat[oref].dev : nullptr;
         if(bucketed)
            hipLaunchKernelGGL(k_item_write_pos, blocks(m), dim3(kB), 0, s, (long long)m, sel, ss, kf, nv, nold, orf,
                               d->ItmLen, (int *)d->dev, (int *)g->dat[g->RefIdx[typ]].dev);
         else
            hipLaunchKernelGGL(k_item_write, blocks(m), dim3(kB), 0, s, (long long)m, u2, p2, kf, it.klo, it.khi,
                               nold, orf, d->ItmLen, (int *)d->dev, (int *)g->dat[g->RefIdx[typ]].dev);
         ok = ok_(hipGetLastError()) && ok_(hipStreamSynchronize(s));
         if(old)
         {
            GmlFreeData(GmlIdx, old);
            if(oref && g->dat[oref].use) FreeDataSlot(g, oref);
         }
         if(ok) rc = m;
      }
   }
   if(!ok) (void)hipGetLastError();
   (void)hipStreamSynchronize(s);
   for(void *p : {(void *)cnt, (void *)nsel, (void *)pos, (void *)first, (void *)hsel, (void *)usel, (void *)p2,
                  (void *)u2, (void *)head, (void *)bcnt, (void *)boff, (void *)slot, (void *)sel, (void *)flag, xtmp,
                  (void *)ccnt, (void *)coff, brec, (void *)defer})
      ItemFree(g, p);
   if(rc < 0 && !hbad) fprintf(stderr, "gml: GPU extraction of %lld items failed\n", total);
   tim[4] = clock();
   DbgPrint(g, "Extraction of %lld items (%lld after the block pass, %d distinct, %s path): allocation %.1f ms, "
            "block pass %.1f ms, first appearances %.1f ms, table and frees %.1f ms\n", total, n2, m,
            viaball ? "vertex-ball" : (bucketed ? "bucketed" : "sorted"), 1e3 * (tim[1] - tim[0]), 1e3 * (tim[2] - tim[1]),
            1e3 * (tim[3] - tim[2]), 1e3 * (tim[4] - tim[3]));
   return rc;
}

// Downlink Src -> Dst (Dst an edge or face kind below Src): the link table of
// LenMatBas[Src][Dst] columns (padded as GetCntVec) per Src element; column
// j = the Dst entity with the nodes of local item j, 0 if none.  Returns the
// data index, 0 on failure.
int GpuDownlink(GmlCtx *g, int Src, int Dst, const std::string &nam)
{
   const int deg = NodesOf(Dst);
   const Data *sd = &g->dat[g->TypIdx[Src]], *dd = &g->dat[g->TypIdx[Dst]];
   const long long nv = g->NmbEle[GmlVertices];
   const KeyForm kf = MakeKeyForm(deg, nv);
   const Source dsrc = {(const int *)dd->dev, dd->NmbLin, dd->ItmLen, Itself(deg)};
   const Source ssrc = {(const int *)sd->dev, sd->NmbLin, sd->ItmLen, ItemsOf(Src, deg)};
   if(dd->NmbLin >= INT_MAX || ssrc.ne * ssrc.li.n >= INT_MAX) return 0;
   int w = LenMatBas[Src][Dst], cnt, vec, typ;
   GetCntVec(w, &cnt, &vec, &typ);
   const int rw = cnt * vec;
   const int idx = NewDatIdx(g);
   if(!idx) return 0;
   Data *d = &g->dat[idx];
   d->kind = DkLnk; d->MshTyp = Src; d->LnkTyp = Dst;
   d->ItmTyp = typ; d->NmbItm = cnt; d->ItmSiz = OclTypSiz[typ]; d->ItmLen = vec;
   d->NmbLin = sd->NmbLin; d->LinSiz = (size_t)rw * 4; d->MemSiz = (size_t)sd->NmbLin * d->LinSiz;
   d->nam = nam;
   if(!AllocData(g, d)) return 0;
   hipStream_t s = g->stream;
   int *bad = nullptr, hbad = 0;
   bool ok = DAlloc(g, &bad, 4) && ok_(hipMemsetAsync(bad, 0, 4, s));
   {
      ItemSort it(g);
      QueryKeys q(g);
      ok = ok && SortItems(g, {dsrc}, kf, nv, bad, &it) && KeysOf(g, ssrc, kf, nv, bad, &q);
      if(ok && q.n && it.n)
         hipLaunchKernelGGL(k_down_lookup, blocks(q.n), dim3(kB), 0, s, q.n, ssrc.li.n, ssrc.li, q.lo, q.hi, it.n,
                            it.slo, it.shi, it.ord, rw, (int *)d->dev);
      ok = ok && ok_(hipGetLastError()) && ok_(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, s))
        && ok_(hipStreamSynchronize(s));
   }
   ItemFree(g, bad);
   if(!ok || hbad)
   {
      (void)hipGetLastError();
      fprintf(stderr, "gml: downlink %s -> %s failed%s\n", BalTypStr[Src], BalTypStr[Dst],
              hbad ? ": a node index is outside the vertex table" : "");
      FreeDataSlot(g, idx);
      return 0;
   }
   d->host_valid = false;
   BumpVersion(d);
   return idx;
}

// Shell of each Src entity (an edge or face) over the Dst elements: degrees
// and SELL-64 rows in descending element order, (element << 4) | local
// index.  Returns 1 on success with the device arrays owned by the caller.
int GpuShell(GmlCtx *g, int Src, int Dst, int *degdev, unsigned int **off, int **ent, long long *nslab,
             int *maxdeg, int *firsthigh)
{
   const int deg = NodesOf(Src);
   const Data *sd = &g->dat[g->TypIdx[Src]], *dd = &g->dat[g->TypIdx[Dst]];
   const long long nv = g->NmbEle[GmlVertices];
   const KeyForm kf = MakeKeyForm(deg, nv);
   const Source dsrc = {(const int *)dd->dev, dd->NmbLin, dd->ItmLen, ItemsOf(Dst, deg)};
   const Source ssrc = {(const int *)sd->dev, sd->NmbLin, sd->ItmLen, Itself(deg)};
   if(dsrc.ne * dsrc.li.n >= INT_MAX || ssrc.ne >= INT_MAX || !dsrc.li.n) return 0;
   hipStream_t s = g->stream;
   int *bad = nullptr, *beg = nullptr, hbad = 0;
   *off = nullptr;
   *ent = nullptr;
   bool ok = DAlloc(g, &bad, 4) && ok_(hipMemsetAsync(bad, 0, 4, s)) && DAlloc(g, &beg, ssrc.ne * 4 + 4)
          && ok_(hipMemsetAsync(degdev, 0, ssrc.ne * 4 + 4, s));
   {
      ItemSort it(g);
      QueryKeys q(g);
      ok = ok && SortItems(g, {dsrc}, kf, nv, bad, &it) && KeysOf(g, ssrc, kf, nv, bad, &q);
      if(ok && q.n && it.n)
         hipLaunchKernelGGL(k_shell_runs, blocks(q.n), dim3(kB), 0, s, q.n, q.lo, q.hi, it.n, it.slo, it.shi, beg,
                            degdev);
      ok = ok && ok_(hipGetLastError()) && ok_(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, s))
        && ok_(hipStreamSynchronize(s)) && !hbad
        && gpu_sell_layout(s, degdev, ssrc.ne, LenMatBas[Src][Dst], off, ent, nslab, maxdeg, firsthigh);
      if(ok && q.n)
         hipLaunchKernelGGL(k_shell_fill, blocks(q.n), dim3(kB), 0, s, q.n, beg, degdev, it.ord, dsrc.li.n, dsrc.li,
                            *off, *ent);
      ok = ok && ok_(hipGetLastError()) && ok_(hipStreamSynchronize(s));
   }
   ItemFree(g, bad);
   ItemFree(g, beg);
   if(!ok)
   {
      (void)hipGetLastError();
      if(*off) (void)hipFree(*off);
      if(*ent) (void)hipFree(*ent);
      *off = nullptr;
      *ent = nullptr;
      fprintf(stderr, "gml: shell %s -> %s failed%s\n", BalTypStr[Src], BalTypStr[Dst],
              hbad ? ": a node index is outside the vertex table" : "");
      return 0;
   }
   return 1;
}

// Neighbours of the elements of kind typ across their items of kind ngb
// into the link data idx (rw ints per element, zeroed).  Returns 1 on
// success.
int GpuNeighbours(GmlCtx *g, int typ, int ngb, int idx, int rw)
{
   const int deg = NodesOf(ngb);
   const Data *td = &g->dat[g->TypIdx[typ]];
   const long long nv = g->NmbEle[GmlVertices];
   const KeyForm kf = MakeKeyForm(deg, nv);
   const Source src = {(const int *)td->dev, td->NmbLin, td->ItmLen, ItemsOf(typ, deg)};
   if(src.ne * src.li.n >= INT_MAX) return 0;
   hipStream_t s = g->stream;
   int *bad = nullptr, hbad = 0;
   bool ok = DAlloc(g, &bad, 4) && ok_(hipMemsetAsync(bad, 0, 4, s));
   {
      ItemSort it(g);
      ok = ok && SortItems(g, {src}, kf, nv, bad, &it);
      if(ok && it.n)
         hipLaunchKernelGGL(k_neighbours, blocks(it.n), dim3(kB), 0, s, it.n, it.slo, it.shi, it.ord, src.li.n, src.li,
                            rw, (int *)g->dat[idx].dev);
      ok = ok && ok_(hipGetLastError()) && ok_(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, s))
        && ok_(hipStreamSynchronize(s));
   }
   ItemFree(g, bad);
   if(!ok || hbad)
   {
      (void)hipGetLastError();
      fprintf(stderr, "gml: neighbours of %s failed%s\n", BalTypStr[typ],
              hbad ? ": a node index is outside the vertex table" : "");
      return 0;
   }
   g->dat[idx].host_valid = false;
   BumpVersion(&g->dat[idx]);
   return 1;
}

} // namespace gml
