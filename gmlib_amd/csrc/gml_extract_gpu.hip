// GPU form of the edge and face extraction (GmlExtractEdges /
// GmlExtractFaces on a device handle).
//
// The reference hashes every item of every element on one CPU thread
// (sources/gmlib3.c:3719-3873 edges, 3880-4171 faces).  Here the items are
// laid out in the reference's visiting order -- the existing table's items,
// then element kind by kind, element by element, local item by local item --
// so that an item's index IS its visiting rank, and each gets a key: its
// sorted nodes packed into one 64-bit word (b bits per node, b from the
// vertex count) or, when deg * b > 64, into two words (32 bits per node).
//   1. Stable LSD radix sort of the item indices by key: equal items become
//      adjacent, each run in visiting order, so its head is the item's
//      first appearance.
//   2. Select the heads, sort them by index: the reference's output order.
//   3. Write the new table from the keys (nodes sorted, as the reference
//      stores them) and the refs (an existing item keeps its ref, others 0).
// All work stays on the device; one 4-byte readback gives the item count.

#include "gml_internal.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <vector>

namespace gml {

namespace {

// the items of one element kind that have the wanted degree, in local order
struct LocalItems {
   int n;
   int nod[12][4];
};

struct Source {
   const int *ele;
   long long ne;
   int len;
   LocalItems li;
};

struct KeyForm {
   int deg, b;        // nodes per item, bits per node in the one-word form
   bool two;          // two-word form: 32 bits per node, hi = leading nodes
   int lo_bits, hi_bits;
};

__device__ __forceinline__ void sort_nodes(int *a, int k)
{
   for(int i = 1; i < k; i++)
      for(int j = i; j > 0 && a[j - 1] > a[j]; j--) { const int t = a[j]; a[j] = a[j - 1]; a[j - 1] = t; }
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
   int a[4] = {0, 0, 0, 0};
   bool out = false;
   for(int k = 0; k < kf.deg; k++)
   {
      a[k] = ele[e * len + li.nod[j][k]];
      out |= a[k] < 0 || a[k] >= nv;
   }
   if(out) *bad = 1;
   sort_nodes(a, kf.deg);
   if(!kf.two)
   {
      unsigned long long k = 0;
      for(int i = 0; i < kf.deg; i++) k = (k << kf.b) | (unsigned)a[i];
      klo[t] = k;
   }
   else
   {
      const unsigned long long a0 = (unsigned)a[0], a1 = (unsigned)a[1], a2 = (unsigned)a[2], a3 = (unsigned)a[3];
      if(kf.deg == 3) { khi[t] = a0; klo[t] = (a1 << 32) | a2; }
      else            { khi[t] = (a0 << 32) | a1; klo[t] = (a2 << 32) | a3; }
   }
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

// heads of the runs of equal keys in sorted order.  One-word form: the
// sorted keys skey; two-word form: sorted high words shi, low words read
// through the order.
__global__ void k_item_heads(long long n, const unsigned long long *__restrict__ skey,
                             const unsigned long long *__restrict__ shi, const int *__restrict__ ord,
                             const unsigned long long *__restrict__ klo, unsigned char *__restrict__ head)
{
   const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(i >= n) return;
   bool h = i == 0;
   if(!h)
   {
      if(!shi) h = skey[i] != skey[i - 1];
      else     h = shi[i] != shi[i - 1] || klo[ord[i]] != klo[ord[i - 1]];
   }
   head[i] = h;
}

// row k of the new table: the nodes of item u[k] decoded from its key,
// its ref (the existing table's ref for an existing item, else 0)
__global__ void k_item_write(long long m, const int *__restrict__ u, KeyForm kf,
                             const unsigned long long *__restrict__ klo, const unsigned long long *__restrict__ khi,
                             long long nold, const int *__restrict__ oref, int w, int *__restrict__ tab,
                             int *__restrict__ ref)
{
   const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(k >= m) return;
   const int i = u[k];
   int a[4] = {0, 0, 0, 0};
   if(!kf.two)
   {
      unsigned long long key = klo[i];
      const unsigned long long mask = (1ull << kf.b) - 1;
      for(int j = kf.deg - 1; j >= 0; j--) { a[j] = (int)(key & mask); key >>= kf.b; }
   }
   else
   {
      const unsigned long long lo = klo[i], hi = khi[i];
      if(kf.deg == 3) { a[0] = (int)hi; a[1] = (int)(lo >> 32); a[2] = (int)(unsigned)lo; }
      else            { a[0] = (int)(hi >> 32); a[1] = (int)(unsigned)hi; a[2] = (int)(lo >> 32); a[3] = (int)(unsigned)lo; }
   }
   for(int j = 0; j < w; j++) tab[k * w + j] = j < kf.deg ? a[j] : 0;
   ref[k] = (i < nold && oref) ? oref[i] : 0;
}

int bits_for(long long n)   // bits to hold 0 .. n-1 (at least 1)
{
   int b = 1;
   while(b < 63 && (1ll << b) < n) b++;
   return b;
}

} // namespace

// Extract the distinct items of degree deg (2: edges, 3: triangles, 4:
// quads) into the table of kind typ, replacing it.  Returns the number of
// items, or -1 on failure (message printed).
int GpuExtractItems(size_t GmlIdx, GmlCtx *g, int typ, int deg)
{
   std::vector<Source> src;
   long long total = 0, nold = 0;
   const int old = g->TypIdx[typ], oref = g->RefIdx[typ];
   if(old)
   {
      const Data *d = &g->dat[old];
      Source s = {(const int *)d->dev, d->NmbLin, d->ItmLen, {1, {{0, 1, 2, 3}}}};
      src.push_back(s);
      nold = d->NmbLin;
   }
   const int t0 = deg == 2 ? GmlEdges + 1 : GmlTetrahedra, t1 = deg == 2 ? GmlMaxEleTyp - 1 : GmlHexahedra;
   for(int t = t0; t <= t1; t++)
   {
      if(!g->TypIdx[t] || !g->NmbEle[t]) continue;
      const Data *d = &g->dat[g->TypIdx[t]];
      Source s = {(const int *)d->dev, d->NmbLin, d->ItmLen, {0, {}}};
      if(deg == 2)
         for(int j = 0; j < NmbTpoLnk[t][GmlEdges]; j++, s.li.n++)
            for(int k = 0; k < 2; k++) s.li.nod[s.li.n][k] = ItmEdgNod[t][j][k];
      else
         for(int j = 0; j < 6 && ItmFacDeg[t][j]; j++)
            if(ItmFacDeg[t][j] == deg)
            {
               for(int k = 0; k < 4; k++) s.li.nod[s.li.n][k] = ItmFacNod[t][j][k];
               s.li.n++;
            }
      if(s.li.n) src.push_back(s);
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

   KeyForm kf;
   kf.deg = deg;
   kf.b = bits_for(std::max(2ll, g->NmbEle[GmlVertices]));
   kf.two = deg * kf.b > 64;
   kf.lo_bits = kf.two ? 32 + kf.b : deg * kf.b;
   kf.hi_bits = kf.two ? (deg == 3 ? kf.b : 32 + kf.b) : 0;

   hipStream_t s = g->stream;
   const unsigned B = 256;
   auto blocks = [&](long long n) { return dim3((unsigned)((n + B - 1) / B)); };
   unsigned long long *klo = nullptr, *khi = nullptr, *ks = nullptr, *ks2 = nullptr;
   int *idx = nullptr, *ord = nullptr, *uni = nullptr, *uni2 = nullptr, *cnt = nullptr;
   unsigned char *head = nullptr;
   void *tmp = nullptr;
   size_t tb[4] = {0, 0, 0, 0}, tbytes = 0;
   const int ibits = bits_for(total);
   int rc = -1, hcnt[2] = {0, 0};
   auto ok_ = [](hipError_t e) { return e == hipSuccess; };
   bool ok = ok_(hipMalloc(&klo, total * 8)) && ok_(hipMalloc(&ks, total * 8))
          && (!kf.two || (ok_(hipMalloc(&khi, total * 8)) && ok_(hipMalloc(&ks2, total * 8))))
          && ok_(hipMalloc(&idx, total * 4)) && ok_(hipMalloc(&ord, total * 4))
          && ok_(hipMalloc(&uni, total * 4)) && ok_(hipMalloc(&uni2, total * 4))
          && ok_(hipMalloc(&head, total)) && ok_(hipMalloc(&cnt, 8));
   if(ok)
      ok = ok_(hipcub::DeviceRadixSort::SortPairs(nullptr, tb[0], klo, ks, idx, ord, (int)total, 0, 64, s))
        && ok_(hipcub::DeviceSelect::Flagged(nullptr, tb[1], ord, head, uni, cnt, (int64_t)total, s))
        && ok_(hipcub::DeviceRadixSort::SortKeys(nullptr, tb[2], uni, uni2, (int)total, 0, 32, s));
   if(ok)
   {
      tbytes = std::max(tb[0], std::max(tb[1], tb[2])) + 256;
      ok = ok_(hipMalloc(&tmp, tbytes)) && ok_(hipMemsetAsync(cnt, 0, 8, s));
   }
   // 1. keys, in visiting order
   long long at = 0;
   for(const Source &sr : src)
   {
      if(!ok) break;
      const long long n = sr.ne * sr.li.n;
      hipLaunchKernelGGL(k_item_keys, blocks(n), dim3(B), 0, s, sr.ne, sr.len, sr.ele, sr.li, kf,
                         g->NmbEle[GmlVertices], klo + at, kf.two ? khi + at : nullptr, cnt + 1);
      at += n;
   }
   // 2. stable LSD sort of the indices: low word, then high word
   if(ok)
   {
      hipLaunchKernelGGL(k_iota, blocks(total), dim3(B), 0, s, total, idx);
      size_t t = tbytes;
      ok = ok_(hipcub::DeviceRadixSort::SortPairs(tmp, t, klo, ks, idx, ord, (int)total, 0, kf.lo_bits, s));
   }
   if(ok && kf.two)
   {
      hipLaunchKernelGGL(k_gather_u64, blocks(total), dim3(B), 0, s, total, ord, khi, ks2);
      size_t t = tbytes;
      ok = ok_(hipcub::DeviceRadixSort::SortPairs(tmp, t, ks2, ks, ord, idx, (int)total, 0, kf.hi_bits, s));
      std::swap(idx, ord);   // ord: final order, ks: sorted high words
   }
   // 3. first appearances, in visiting order
   if(ok)
   {
      hipLaunchKernelGGL(k_item_heads, blocks(total), dim3(B), 0, s, total, kf.two ? nullptr : ks,
                         kf.two ? ks : nullptr, ord, klo, head);
      size_t t = tbytes;
      ok = ok_(hipcub::DeviceSelect::Flagged(tmp, t, ord, head, uni, cnt, (int64_t)total, s))
        && ok_(hipMemcpyAsync(hcnt, cnt, 8, hipMemcpyDeviceToHost, s)) && ok_(hipStreamSynchronize(s));
   }
   if(ok && hcnt[1])
   {
      fprintf(stderr, "gml: extraction: an element references a vertex outside 0..%lld\n", g->NmbEle[GmlVertices] - 1);
      ok = false;
      rc = -1;
   }
   const long long m = hcnt[0];
   if(ok)
   {
      size_t t = tbytes;
      ok = ok_(hipcub::DeviceRadixSort::SortKeys(tmp, t, uni, uni2, (int)m, 0, ibits, s));
   }
   // 4. the new table (created before the old one is dropped: the old refs are read)
   if(ok)
   {
      const int nidx = GmlNewMeshData(GmlIdx, typ, (int)m);
      if(!nidx) ok = false;
      else
      {
         const Data *d = &g->dat[nidx];
         hipLaunchKernelGGL(k_item_write, blocks(m), dim3(B), 0, s, m, uni2, kf, klo, khi, nold,
                            oref ? (const int *)g->dat[oref].dev : nullptr, d->ItmLen, (int *)d->dev,
                            (int *)g->dat[g->RefIdx[typ]].dev);
         ok = ok_(hipGetLastError()) && ok_(hipStreamSynchronize(s));
         if(old)
         {
            GmlFreeData(GmlIdx, old);
            if(oref && g->dat[oref].use) FreeDataSlot(g, oref);
         }
         if(ok) rc = (int)m;
      }
   }
   if(!ok) (void)hipGetLastError();
   (void)hipStreamSynchronize(s);
   for(void *p : {(void *)klo, (void *)khi, (void *)ks, (void *)ks2, (void *)idx, (void *)ord, (void *)uni,
                  (void *)uni2, (void *)head, (void *)cnt, tmp})
      if(p) (void)hipFree(p);
   if(rc < 0 && !hcnt[1]) fprintf(stderr, "gml: GPU extraction of %lld items failed\n", total);
   return rc;
}

} // namespace gml
