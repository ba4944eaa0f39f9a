// Edge-graph vector Laplacian of a tet mesh (host), the matrix of the block
// CG jobs: 3x3 blocks, block row per vertex, columns ascending (the diagonal
// at its sorted place),
//    off-diagonal block (i, j) = -M_e,   M_e = w_e (I + 0.25 s_e s_e^T),
//    diagonal block (i, i)     = sum of M_e over the edges at i + shift I,
// w_e in [0.5, 2] and s_e a unit vector drawn from a counter-based hash of
// (seed, i, j), i < j, so the matrix depends only on the mesh and the seed.
// Sum of PSD edge terms plus shift I: symmetric positive definite.  Entries
// are formed in FP64 and stored in FP32.

#include "gml_internal.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>

namespace gml {

static inline unsigned long long Mix64(unsigned long long z)
{
   z += 0x9e3779b97f4a7c15ULL;
   z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
   z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
   return z ^ (z >> 31);
}

// uniform in [0, 1) from the top 53 bits
static inline double Unit(unsigned long long z) { return (double)(z >> 11) * (1.0 / 9007199254740992.0); }

// M_e of edge (i, j), i < j: 9 doubles, row-major
static void EdgeBlock(unsigned long long seed, long long i, long long j, double *M)
{
   const unsigned long long k = Mix64(seed ^ Mix64(((unsigned long long)i << 32) | (unsigned long long)j));
   const double w = 0.5 + 1.5 * Unit(Mix64(k + 1));
   double s[3];
   for(int c = 0; c < 3; c++) s[c] = 2.0 * Unit(Mix64(k + 2 + c)) - 1.0;
   double n = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
   if(n < 1e-3) { s[0] = 1.0; s[1] = 0.0; s[2] = 0.0; n = 1.0; }
   for(int c = 0; c < 3; c++) s[c] /= n;
   for(int a = 0; a < 3; a++)
      for(int b = 0; b < 3; b++) M[a * 3 + b] = w * ((a == b ? 1.0 : 0.0) + 0.25 * s[a] * s[b]);
}

} // namespace gml

using namespace gml;

extern "C" {

long long GmlxEdgeLaplacian3(int NmbVer, int NmbTet, const int *tet, unsigned long long seed, double shift,
                             int *rowptr, int *col, float *val, long long cap)
{
   if(NmbVer < 1 || NmbTet < 0 || !tet || !rowptr) return -1;
   for(long long e = 0; e < (long long)NmbTet * 4; e++)
      if(tet[e] < 0 || tet[e] >= NmbVer) return -1;
   const long long V = NmbVer, T = NmbTet;
   // vertex -> tet incidence (counting sort)
   std::vector<long long> bo((size_t)V + 1, 0);
   for(long long e = 0; e < T * 4; e++) bo[tet[e] + 1]++;
   for(long long v = 0; v < V; v++) bo[v + 1] += bo[v];
   std::vector<int> ball((size_t)bo[V]);
   {
      std::vector<long long> f(bo.begin(), bo.end() - 1);
      for(long long e = 0; e < T; e++)
         for(int j = 0; j < 4; j++) ball[(size_t)f[tet[e * 4 + j]]++] = (int)e;
   }
   const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
   auto parallel = [&](auto body) {
      std::vector<std::thread> th;
      for(unsigned w = 0; w < nth; w++)
         th.emplace_back([&, w]() {
            std::vector<int> nb;
            for(long long v = w; v < V; v += nth) body(v, nb);
         });
      for(auto &x : th) x.join();
   };
   // neighbours of v (itself included), ascending
   auto neighbours = [&](long long v, std::vector<int> &nb) {
      nb.clear();
      for(long long k = bo[v]; k < bo[v + 1]; k++)
         for(int j = 0; j < 4; j++) nb.push_back(tet[(size_t)ball[k] * 4 + j]);
      if(bo[v] == bo[v + 1]) nb.push_back((int)v);
      std::sort(nb.begin(), nb.end());
      nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
   };
   std::vector<int> cnt((size_t)V);
   parallel([&](long long v, std::vector<int> &nb) { neighbours(v, nb); cnt[v] = (int)nb.size(); });
   long long nnz = 0;
   rowptr[0] = 0;
   for(long long v = 0; v < V; v++)
   {
      nnz += cnt[v];
      if(nnz > 0x7fffffffLL) return -2;
      rowptr[v + 1] = (int)nnz;
   }
   if(!col || !val || cap < nnz) return nnz;
   parallel([&](long long v, std::vector<int> &nb) {
      neighbours(v, nb);
      double D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, M[9];
      long long di = -1;
      for(size_t k = 0; k < nb.size(); k++)
      {
         const long long u = nb[k], p = rowptr[v] + (long long)k;
         col[p] = (int)u;
         if(u == v) { di = p; continue; }
         EdgeBlock(seed, std::min(u, v), std::max(u, v), M);
         for(int c = 0; c < 9; c++)
         {
            val[p * 9 + c] = (float)(-M[c]);
            D[c] += M[c];
         }
      }
      for(int c = 0; c < 9; c += 4) D[c] += shift;
      for(int c = 0; c < 9; c++) val[di * 9 + c] = (float)D[c];
   });
   return nnz;
}

} // extern "C"
