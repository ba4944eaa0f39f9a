// Synthetic cube tet meshes for the workloads of this engine: n^3 cells of
// the unit cube, each split into the 6 Kuhn tets that share its main
// diagonal (a conforming split: every cell uses the same diagonal
// direction), with seeded jitter and optional 3D Hilbert renumbering of
// vertices and cells.
//
// Orientation: every tet has a positive CalVol (sources/toolkit.cl:84-87),
// dot(cross(b-a, c-a), d-a) > 0 before jitter.
// Hilbert keys: J. Skilling, "Programming the Hilbert curve", AIP Conf.
// Proc. 707 (2004) -- axes-to-transpose transform, then bit interleave.

#include "../../include/gmlx.h"
#include "gml_hilbert.h"

#include <algorithm>
#include <cstdint>
#include <thread>
#include <vector>

namespace {

using gml::HilbertKey;

uint64_t SplitMix(uint64_t x)
{
   x += 0x9e3779b97f4a7c15ull;
   x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
   x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
   return x ^ (x >> 31);
}

// uniform in [-1, 1) from (seed, vertex, axis)
double Jit(uint64_t seed, uint64_t v, int axis)
{
   uint64_t h = SplitMix(seed ^ SplitMix(v * 3 + (uint64_t)axis + 0x51ed270b27a8a4d5ull));
   return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

template<class F> void ParFor(long long n, F f)
{
   unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
   if(n < 65536) nt = 1;
   std::vector<std::thread> th;
   for(unsigned t = 0; t < nt; t++)
      th.emplace_back([=]() {
         long long b = n * t / nt, e = n * (t + 1) / nt;
         for(long long i = b; i < e; i++) f(i);
      });
   for(auto &x : th) x.join();
}

// permutation order of the 6 Kuhn tets of a cell (axis sequence)
const int kPerm[6][3] = {{0,1,2}, {0,2,1}, {1,0,2}, {1,2,0}, {2,0,1}, {2,1,0}};
const int kSign[6] = {+1, -1, -1, +1, +1, -1};

} // namespace

extern "C" int GmlxCubeMesh(int n, double jitter, unsigned long long seed, int order,
                            float *crd, int *ref, int *tet)
{
   if(n < 1 || n > 1000 || !crd || !ref || !tet) return 0;
   const long long nv1 = n + 1, nver = nv1 * nv1 * nv1, ncell = (long long)n * n * n;
   const double h = 1.0 / n;
   int bits = 1;
   while((1LL << bits) < nv1) bits++;

   // vertex numbering
   std::vector<int> vnew((size_t)nver), cord((size_t)ncell);
   if(order)
   {
      std::vector<std::pair<uint64_t, int>> k((size_t)nver);
      ParFor(nver, [&](long long v) {
         unsigned i = (unsigned)(v % nv1), j = (unsigned)((v / nv1) % nv1), l = (unsigned)(v / (nv1 * nv1));
         k[v] = {HilbertKey(i, j, l, bits), (int)v};
      });
      std::sort(k.begin(), k.end());
      for(long long r = 0; r < nver; r++) vnew[k[r].second] = (int)r;
      std::vector<std::pair<uint64_t, int>> c((size_t)ncell);
      ParFor(ncell, [&](long long q) {
         unsigned i = (unsigned)(q % n), j = (unsigned)((q / n) % n), l = (unsigned)(q / ((long long)n * n));
         c[q] = {HilbertKey(i, j, l, bits), (int)q};
      });
      std::sort(c.begin(), c.end());
      for(long long r = 0; r < ncell; r++) cord[r] = c[r].second;
   }
   else
   {
      for(long long v = 0; v < nver; v++) vnew[v] = (int)v;
      for(long long q = 0; q < ncell; q++) cord[q] = (int)q;
   }

   // coordinates, refs
   ParFor(nver, [&](long long v) {
      long long id[3] = {v % nv1, (v / nv1) % nv1, v / (nv1 * nv1)};
      int bnd = 0;
      float *c = &crd[(size_t)vnew[v] * 4];
      for(int a = 0; a < 3; a++)
      {
         double x = id[a] * h;
         if(id[a] == 0 || id[a] == n) bnd = 1;
         else if(jitter != 0.) x += jitter * h * Jit(seed, (uint64_t)v, a);
         c[a] = (float)x;
      }
      c[3] = 0.f;
      ref[vnew[v]] = bnd;
   });

   // tets, 6 per cell, cells in the chosen order
   ParFor(ncell, [&](long long r) {
      long long q = cord[r];
      long long i = q % n, j = (q / n) % n, l = q / ((long long)n * n);
      for(int p = 0; p < 6; p++)
      {
         long long c[3] = {i, j, l}, vid[4];
         for(int s = 0; s < 4; s++)
         {
            vid[s] = c[0] + nv1 * (c[1] + nv1 * c[2]);
            if(s < 3) c[kPerm[p][s]]++;
         }
         if(kSign[p] < 0) std::swap(vid[2], vid[3]);
         int *t = &tet[((size_t)r * 6 + p) * 4];
         for(int s = 0; s < 4; s++) t[s] = vnew[vid[s]];
      }
   });
   return 1;
}
