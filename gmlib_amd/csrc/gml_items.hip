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
// sorted nodes packed into one 64-bit word (b bits per node, b from the
// vertex count) or, when deg * b > 64, into two words (32 bits per node).
// A stable LSD radix sort of the item indices by key puts equal items next
// to each other, each run in visiting order (ascending element), and every
// query becomes a run lookup:
//   extraction  the runs' heads, selected and sorted by index;
//   downlink    the last item of the run (the highest entity index, the
//               head of the reference's chain);
//   shell       the run reversed (descending element, the chain order);
//   neighbour   the lowest element of the run other than the element itself.

#include "gml_internal.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
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
   int deg, b;        // nodes per item, bits per node in the one-word form
   bool two;          // two-word form: 32 bits per node, hi = leading nodes
   int lo_bits, hi_bits;
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
   kf.lo_bits = kf.two ? 32 + kf.b : deg * kf.b;
   kf.hi_bits = kf.two ? (deg == 3 ? kf.b : 32 + kf.b) : 0;
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
   long long n = 0;
   KeyForm kf;
   unsigned long long *klo = nullptr, *khi = nullptr, *slo = nullptr, *shi = nullptr;
   int *ord = nullptr, *idx = nullptr;
   void *tmp = nullptr;
   size_t tbytes = 0;
   ~ItemSort()
   {
      for(void *p : {(void *)klo, (void *)khi, (void *)slo, (void *)shi, (void *)ord, (void *)idx, tmp})
         if(p) (void)hipFree(p);
   }
};

const unsigned kB = 256;
dim3 blocks(long long n) { return dim3((unsigned)((n + kB - 1) / kB)); }
bool ok_(hipError_t e) { return e == hipSuccess; }

// Keys of every item of the sources (in order) and their stable sort.
// bad: device flag raised by a node out of range.  The temporary storage
// also fits a select and a 32-bit key sort over n items.
bool SortItems(hipStream_t s, const std::vector<Source> &src, const KeyForm &kf, long long nv, int *bad,
               ItemSort *it)
{
   long long n = 0;
   for(const Source &sr : src) n += sr.ne * sr.li.n;
   it->n = n;
   it->kf = kf;
   if(!n) return true;
   size_t t1 = 0, t2 = 0, t3 = 0;
   int *cnt = nullptr;
   unsigned char *flags = nullptr;
   bool ok = ok_(hipMalloc(&it->klo, n * 8)) && ok_(hipMalloc(&it->slo, n * 8))
          && (!kf.two || (ok_(hipMalloc(&it->khi, n * 8)) && ok_(hipMalloc(&it->shi, n * 8))))
          && ok_(hipMalloc(&it->ord, n * 4)) && ok_(hipMalloc(&it->idx, n * 4))
          && ok_(hipcub::DeviceRadixSort::SortPairs(nullptr, t1, it->klo, it->slo, it->idx, it->ord, (int)n, 0, 64, s))
          && ok_(hipcub::DeviceSelect::Flagged(nullptr, t2, it->ord, flags, it->idx, cnt, (int64_t)n, s))
          && ok_(hipcub::DeviceRadixSort::SortKeys(nullptr, t3, it->ord, it->idx, (int)n, 0, 32, s));
   if(ok)
   {
      it->tbytes = std::max(t1, std::max(t2, t3)) + 256;
      ok = ok_(hipMalloc(&it->tmp, it->tbytes));
   }
   long long at = 0;
   for(const Source &sr : src)
   {
      if(!ok) break;
      const long long m = sr.ne * sr.li.n;
      if(m)
         hipLaunchKernelGGL(k_item_keys, blocks(m), dim3(kB), 0, s, sr.ne, sr.len, sr.ele, sr.li, kf, nv,
                            it->klo + at, kf.two ? it->khi + at : nullptr, bad);
      at += m;
   }
   if(ok)
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

// queries: keys of the items of the sources, unsorted
struct QueryKeys {
   long long n = 0;
   unsigned long long *lo = nullptr, *hi = nullptr;
   ~QueryKeys()
   {
      if(lo) (void)hipFree(lo);
      if(hi) (void)hipFree(hi);
   }
};

bool KeysOf(hipStream_t s, const Source &sr, const KeyForm &kf, long long nv, int *bad, QueryKeys *q)
{
   q->n = sr.ne * sr.li.n;
   if(!q->n) return true;
   if(!ok_(hipMalloc(&q->lo, q->n * 8)) || (kf.two && !ok_(hipMalloc(&q->hi, q->n * 8)))) return false;
   hipLaunchKernelGGL(k_item_keys, blocks(q->n), dim3(kB), 0, s, sr.ne, sr.len, sr.ele, sr.li, kf, nv, q->lo, q->hi,
                      bad);
   return ok_(hipGetLastError());
}

// ---------------------------------------------------------------------------
// extraction
// ---------------------------------------------------------------------------

// heads of the runs of equal keys in sorted order
__global__ void k_item_heads(long long n, const unsigned long long *__restrict__ slo,
                             const unsigned long long *__restrict__ shi, unsigned char *__restrict__ head)
{
   const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(i >= n) return;
   head[i] = i == 0 || slo[i] != slo[i - 1] || (shi && shi[i] != shi[i - 1]);
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
   long long total = 0, nold = 0;
   const int old = g->TypIdx[typ], oref = g->RefIdx[typ];
   if(old)
   {
      const Data *d = &g->dat[old];
      src.push_back({(const int *)d->dev, d->NmbLin, d->ItmLen, Itself(deg)});
      nold = d->NmbLin;
   }
   const int t0 = deg == 2 ? GmlEdges + 1 : GmlTetrahedra, t1 = deg == 2 ? GmlMaxEleTyp - 1 : GmlHexahedra;
   for(int t = t0; t <= t1; t++)
   {
      if(!g->TypIdx[t] || !g->NmbEle[t]) continue;
      const Data *d = &g->dat[g->TypIdx[t]];
      const Source s = {(const int *)d->dev, d->NmbLin, d->ItmLen, ItemsOf(t, deg)};
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

   hipStream_t s = g->stream;
   const long long nv = g->NmbEle[GmlVertices];
   const KeyForm kf = MakeKeyForm(deg, nv);
   ItemSort it;
   int *cnt = nullptr, *uni = nullptr, *uni2 = nullptr;
   unsigned char *head = nullptr;
   int rc = -1, hcnt[2] = {0, 0};
   bool ok = ok_(hipMalloc(&cnt, 8)) && ok_(hipMemsetAsync(cnt, 0, 8, s))
          && SortItems(s, src, kf, nv, cnt + 1, &it)
          && ok_(hipMalloc(&uni, total * 4)) && ok_(hipMalloc(&uni2, total * 4)) && ok_(hipMalloc(&head, total));
   // first appearances, in visiting order
   if(ok)
   {
      hipLaunchKernelGGL(k_item_heads, blocks(total), dim3(kB), 0, s, total, it.slo, it.shi, head);
      size_t t = it.tbytes;
      ok = ok_(hipcub::DeviceSelect::Flagged(it.tmp, t, it.ord, head, uni, cnt, (int64_t)total, s))
        && ok_(hipMemcpyAsync(hcnt, cnt, 8, hipMemcpyDeviceToHost, s)) && ok_(hipStreamSynchronize(s));
   }
   if(ok && hcnt[1])
   {
      fprintf(stderr, "gml: extraction: an element references a vertex outside 0..%lld\n", nv - 1);
      ok = false;
   }
   const long long m = hcnt[0];
   if(ok)
   {
      size_t t = it.tbytes;
      ok = ok_(hipcub::DeviceRadixSort::SortKeys(it.tmp, t, uni, uni2, (int)m, 0, bits_for(total), s));
   }
   // the new table (created before the old one is dropped: the old refs are read)
   if(ok)
   {
      const int nidx = GmlNewMeshData(GmlIdx, typ, (int)m);
      if(!nidx) ok = false;
      else
      {
         const Data *d = &g->dat[nidx];
         hipLaunchKernelGGL(k_item_write, blocks(m), dim3(kB), 0, s, m, uni2, kf, it.klo, it.khi, nold,
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
   for(void *p : {(void *)uni, (void *)uni2, (void *)head, (void *)cnt})
      if(p) (void)hipFree(p);
   if(rc < 0 && !hcnt[1]) fprintf(stderr, "gml: GPU extraction of %lld items failed\n", total);
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
   bool ok = ok_(hipMalloc(&bad, 4)) && ok_(hipMemsetAsync(bad, 0, 4, s));
   {
      ItemSort it;
      QueryKeys q;
      ok = ok && SortItems(s, {dsrc}, kf, nv, bad, &it) && KeysOf(s, ssrc, kf, nv, bad, &q);
      if(ok && q.n && it.n)
         hipLaunchKernelGGL(k_down_lookup, blocks(q.n), dim3(kB), 0, s, q.n, ssrc.li.n, ssrc.li, q.lo, q.hi, it.n,
                            it.slo, it.shi, it.ord, rw, (int *)d->dev);
      ok = ok && ok_(hipGetLastError()) && ok_(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, s))
        && ok_(hipStreamSynchronize(s));
   }
   if(bad) (void)hipFree(bad);
   if(!ok || hbad)
   {
      (void)hipGetLastError();
      fprintf(stderr, "gml: downlink %s -> %s failed%s\n", BalTypStr[Src], BalTypStr[Dst],
              hbad ? ": a node index is outside the vertex table" : "");
      FreeDataSlot(g, idx);
      return 0;
   }
   d->host_valid = false;
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
   bool ok = ok_(hipMalloc(&bad, 4)) && ok_(hipMemsetAsync(bad, 0, 4, s)) && ok_(hipMalloc(&beg, ssrc.ne * 4 + 4))
          && ok_(hipMemsetAsync(degdev, 0, ssrc.ne * 4 + 4, s));
   {
      ItemSort it;
      QueryKeys q;
      ok = ok && SortItems(s, {dsrc}, kf, nv, bad, &it) && KeysOf(s, ssrc, kf, nv, bad, &q);
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
   if(bad) (void)hipFree(bad);
   if(beg) (void)hipFree(beg);
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
   bool ok = ok_(hipMalloc(&bad, 4)) && ok_(hipMemsetAsync(bad, 0, 4, s));
   {
      ItemSort it;
      ok = ok && SortItems(s, {src}, kf, nv, bad, &it);
      if(ok && it.n)
         hipLaunchKernelGGL(k_neighbours, blocks(it.n), dim3(kB), 0, s, it.n, it.slo, it.shi, it.ord, src.li.n, src.li,
                            rw, (int *)g->dat[idx].dev);
      ok = ok && ok_(hipGetLastError()) && ok_(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, s))
        && ok_(hipStreamSynchronize(s));
   }
   if(bad) (void)hipFree(bad);
   if(!ok || hbad)
   {
      (void)hipGetLastError();
      fprintf(stderr, "gml: neighbours of %s failed%s\n", BalTypStr[typ],
              hbad ? ": a node index is outside the vertex table" : "");
      return 0;
   }
   g->dat[idx].host_valid = false;
   return 1;
}

} // namespace gml
