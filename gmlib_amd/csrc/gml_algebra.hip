// Block-sparse matrices and block vectors of the GMlib C API, plus the
// block-Jacobi preconditioned CG solver built on them.
//
// Reference behaviour followed (paths in the GMlib tree):
//   GmlNewMatrix  (CSR of b x b row-major blocks in)  sources/gmlib3.c:1175-1427
//   GmlNewVector                                      sources/gmlib3.c:1434-1560
//   GmlMultMatVec: vec1 = A . vec2 (the kernel writes its 4th argument,
//     bound to vec1, from the 5th, bound to vec2)     gmlib3.c:3339-3404, multmatvec.cl:30-67
//   GmlMultDiagMatVec: vec3 = D . vec2, D = b^2 per row  gmlib3.c:3582-3641, multdiagmatvec.cl:18-41
//   GmlAddVec3: vec4 = vec1 + vec2 + vec3             gmlib3.c:3411-3488, addvec.cl:18-33
//   GmlScaleVec: vec *= scale                         gmlib3.c:3495-3527, scalevec.cl:24-37
//   GmlNormVec: sqrt(sum over rows of |row|_2)        gmlib3.c:3534-3575, normvec.cl:18-36
//   byte/flop accounting of each op                   gmlib3.c:1417-1420, 3484, 3523, 3571, 3637
//
// Layout (this engine's own): matrices are SELL-64 -- 64 consecutive block
// rows per chunk, chunk width = its longest row, column indices and the b*b
// block values stored slot-major and lane-minor, so every load of a wave is
// one coalesced 256-byte (FP32) row of the chunk.  Any block size 1..8 and
// any row-length distribution is accepted (no degree-sorted slices).
// Vectors are packed n x b arrays.

#include "gml_internal.h"

#include <cmath>
#include <cstring>

namespace gml {

struct CgWork;

struct Matrix {
   int NmbLin = 0, NmbBlk = 0, b = 0, flt = GmlFlt;
   long long nslab = 0;
   int *deg = nullptr, *col = nullptr;
   unsigned int *off = nullptr;
   void *val = nullptr;
   // near slots: in every slab the first kn slots of each row hold columns
   // within +-32767 of the row, also stored as 16-bit offsets from it
   // (slab c: kn = off16[c + 1] - off16[c] slots at col16[(off16[c] + k) * 64 + lane])
   unsigned int *off16 = nullptr;
   short *col16 = nullptr;
   long long nslab16 = 0, n16 = 0;   // slabs with 16-bit columns, blocks read through them
   double MemAcc = 0., FltOpp = 0.;
   CgWork *cg = nullptr;        // CG workspace, kept from the first solve on
};

static void FreeCgWork(GmlCtx *g, CgWork *w);

struct Vector {
   int NmbLin = 0, b = 0, flt = GmlFlt;
   void *dev = nullptr;
   size_t bytes = 0;
};

void FreeAlgebra(GmlCtx *g)
{
   for(int i = 0; i <= GmlMaxMat; i++)
      if(g->mat[i])
      {
         FreeCgWork(g, g->mat[i]->cg);
         DevFree(g, g->mat[i]->deg); DevFree(g, g->mat[i]->col);
         DevFree(g, g->mat[i]->off); DevFree(g, g->mat[i]->val);
         DevFree(g, g->mat[i]->off16); DevFree(g, g->mat[i]->col16);
         delete g->mat[i];
         g->mat[i] = nullptr;
      }
   for(int i = 0; i <= GmlMaxVec; i++)
      if(g->vec[i])
      {
         DevFree(g, g->vec[i]->dev);
         delete g->vec[i];
         g->vec[i] = nullptr;
      }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// One block row per lane.  Slots are taken two at a time, both slots' loads
// issued before either product, so a wave keeps ~2(B*B+B+1) loads in flight;
// the products are added in slot order, as a one-at-a-time loop would.
template<int B, class T, class X>
__device__ __forceinline__ void slot_pair(size_t s0, int j0, int j1, int l, const T *__restrict__ val,
                                          const X *__restrict__ x, X *acc)
{
   const size_t s1 = s0 + 1;
   T v0[B * B], v1[B * B];
#pragma unroll
   for(int e = 0; e < B * B; e++)
   {
      v0[e] = __builtin_nontemporal_load(&val[(s0 * (B * B) + e) * 64 + l]);
      v1[e] = __builtin_nontemporal_load(&val[(s1 * (B * B) + e) * 64 + l]);
   }
   X x0[B], x1[B];
#pragma unroll
   for(int m = 0; m < B; m++) { x0[m] = x[(size_t)j0 * B + m]; x1[m] = x[(size_t)j1 * B + m]; }
#pragma unroll
   for(int i = 0; i < B; i++)
   {
      X t = 0;
#pragma unroll
      for(int m = 0; m < B; m++) t += (X)v0[i * B + m] * x0[m];
      acc[i] += t;
   }
#pragma unroll
   for(int i = 0; i < B; i++)
   {
      X t = 0;
#pragma unroll
      for(int m = 0; m < B; m++) t += (X)v1[i * B + m] * x1[m];
      acc[i] += t;
   }
}

// Row r (lane l of its slab, d slots from slab slot base).  The slab's first
// kn slots (kn <= every row's degree in the slab, so the same count in every
// lane) take their columns from the 16-bit offsets at slab slot base16; the
// rest from the 32-bit columns.  At 256^3 (the edge Laplacian of the 100.7M-tet
// cube) 4% of the columns lie beyond +-32767 of their row.
template<int B, class T, class X>
__device__ __forceinline__ void row_product(int r, int l, int d, size_t base, int kn, size_t base16,
                                            const int *__restrict__ col, const short *__restrict__ col16,
                                            const T *__restrict__ val, const X *__restrict__ x, X *acc)
{
   int k = 0;
   for(; k + 1 < kn; k += 2)
   {
      const int j0 = r + (int)__builtin_nontemporal_load(&col16[(base16 + k) * 64 + l]);
      const int j1 = r + (int)__builtin_nontemporal_load(&col16[(base16 + k + 1) * 64 + l]);
      slot_pair<B, T, X>(base + k, j0, j1, l, val, x, acc);
   }
   for(; k + 1 < d; k += 2)
   {
      const size_t s0 = base + k;
      const int j0 = __builtin_nontemporal_load(&col[s0 * 64 + l]), j1 = __builtin_nontemporal_load(&col[(s0 + 1) * 64 + l]);
      slot_pair<B, T, X>(s0, j0, j1, l, val, x, acc);
   }
   if(k < d)
   {
      const size_t s0 = base + k;
      const int j0 = col[s0 * 64 + l];
#pragma unroll
      for(int i = 0; i < B; i++)
      {
         X t = 0;
#pragma unroll
         for(int m = 0; m < B; m++) t += (X)val[(s0 * (B * B) + i * B + m) * 64 + l] * x[(size_t)j0 * B + m];
         acc[i] += t;
      }
   }
}

template<int B, class T, class X>
__global__ void __launch_bounds__(256) k_spmv(int n, const int *__restrict__ deg, const unsigned int *__restrict__ off,
                                              const int *__restrict__ col, const unsigned int *__restrict__ off16,
                                              const short *__restrict__ col16, const T *__restrict__ val,
                                              const X *__restrict__ x, X *__restrict__ y, int remap)
{
   const int r = (int)(xcd_block(remap) * blockDim.x + threadIdx.x);
   if(r >= n) return;
   X acc[B];
#pragma unroll
   for(int i = 0; i < B; i++) acc[i] = 0;
   const unsigned int b16 = off16[r >> 6];
   row_product<B, T, X>(r, r & 63, deg[r], (size_t)off[r >> 6], (int)(off16[(r >> 6) + 1] - b16), (size_t)b16, col, col16,
                        val, x, acc);
#pragma unroll
   for(int i = 0; i < B; i++) y[(size_t)r * B + i] = acc[i];
}

// Elementwise vector kernels: 16 bytes per load and store (4 floats / 2
// doubles).  A work-group takes one contiguous block of kVecLoads x 256
// vectors (thread t: vectors t, t + 256, ...), every load of the block in
// flight before the stores, streamed past the caches (nontemporal); the
// entries past the last whole vector are done by thread 0 of the last
// work-group.  Each entry is read and written by the same thread, so the
// output may be one of the inputs.
typedef float  vf4 __attribute__((ext_vector_type(4)));
typedef double vd2 __attribute__((ext_vector_type(2)));
template<class T> struct Vec16;
template<> struct Vec16<float>  { typedef vf4 t; static const int n = 4; };
template<> struct Vec16<double> { typedef vd2 t; static const int n = 2; };
static const int kVecLoads = 4;

template<class T>
__global__ void __launch_bounds__(256) k_add3(long long n, const T *u, const T *v, const T *w, T *y)
{
   typedef typename Vec16<T>::t V;
   const int W = Vec16<T>::n;
   const long long nv = n / W;
   const long long i0 = (long long)blockIdx.x * (kVecLoads * 256) + threadIdx.x;
   V a[kVecLoads], b[kVecLoads], c[kVecLoads];
#pragma unroll
   for(int k = 0; k < kVecLoads; k++)
   {
      const long long i = i0 + k * 256;
      if(i < nv)
      {
         a[k] = __builtin_nontemporal_load((const V *)u + i);
         b[k] = __builtin_nontemporal_load((const V *)v + i);
         c[k] = __builtin_nontemporal_load((const V *)w + i);
      }
   }
#pragma unroll
   for(int k = 0; k < kVecLoads; k++)
   {
      const long long i = i0 + k * 256;
      if(i < nv)
      {
         V r = a[k] + b[k];
         r += c[k];
         __builtin_nontemporal_store(r, (V *)y + i);
      }
   }
   if(blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
      for(long long k = nv * W; k < n; k++) y[k] = u[k] + v[k] + w[k];
}

template<class T>
__global__ void __launch_bounds__(256) k_scale(long long n, T *u, T s)
{
   typedef typename Vec16<T>::t V;
   const int W = Vec16<T>::n;
   const long long nv = n / W;
   const long long i0 = (long long)blockIdx.x * (kVecLoads * 256) + threadIdx.x;
   V a[kVecLoads];
#pragma unroll
   for(int k = 0; k < kVecLoads; k++)
      if(i0 + k * 256 < nv) a[k] = __builtin_nontemporal_load((const V *)u + i0 + k * 256);
#pragma unroll
   for(int k = 0; k < kVecLoads; k++)
      if(i0 + k * 256 < nv) __builtin_nontemporal_store(a[k] * s, (V *)u + i0 + k * 256);
   if(blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
      for(long long k = nv * W; k < n; k++) u[k] *= s;
}

// grid of k_add3 / k_scale over n entries (at least one work-group, which
// also does the entries past the last whole vector)
template<class T> static unsigned int VecBlocks(long long n)
{
   const long long w = (long long)Vec16<T>::n * 256 * kVecLoads;
   const long long nb = (n + w - 1) / w;
   return (unsigned int)(nb < 1 ? 1 : nb);
}

template<class T>
__global__ void k_rownorm(int n, int b, const T *__restrict__ u, float *__restrict__ out)
{
   int r = blockIdx.x * blockDim.x + threadIdx.x;
   if(r >= n) return;
   T s = 0;
   for(int i = 0; i < b; i++) { T x = u[(size_t)r * b + i]; s += x * x; }
   out[r] = (float)sqrt(s);
}

template<class T>
__global__ void k_diagmul(int n, int b, const T *__restrict__ D, const T *__restrict__ u, T *__restrict__ v)
{
   int r = blockIdx.x * blockDim.x + threadIdx.x;
   if(r >= n) return;
   const T *d = D + (size_t)r * b * b;
   const T *x = u + (size_t)r * b;
   for(int i = 0; i < b; i++)
   {
      T t = 0;
      for(int m = 0; m < b; m++) t += d[i * b + m] * x[m];
      v[(size_t)r * b + i] = t;
   }
}

static unsigned int Blocks(long long n, int b = 256) { return (unsigned int)((n + b - 1) / b); }

// y = A x with A's values of type T and vectors of type X
template<class T, class X>
static int SpMV(const GmlCtx *g, const Matrix *m, const X *x, X *y)
{
   hipStream_t s = g->stream;
   const Shape sh = BuiltinShape(g, GmlxBuiltinSpMV);
   const dim3 nb(Blocks(m->NmbLin, sh.grp)), gs(sh.grp);
   const int *deg = m->deg, *col = m->col;
   const unsigned int *off = m->off, *o16 = m->off16;
   const short *c16 = m->col16;
   const T *val = (const T *)m->val;
   int n = m->NmbLin, rm = sh.remap;
   switch(m->b)
   {
      case 1: hipLaunchKernelGGL((k_spmv<1, T, X>), nb, gs, 0, s, n, deg, off, col, o16, c16, val, x, y, rm); break;
      case 2: hipLaunchKernelGGL((k_spmv<2, T, X>), nb, gs, 0, s, n, deg, off, col, o16, c16, val, x, y, rm); break;
      case 3: hipLaunchKernelGGL((k_spmv<3, T, X>), nb, gs, 0, s, n, deg, off, col, o16, c16, val, x, y, rm); break;
      case 4: hipLaunchKernelGGL((k_spmv<4, T, X>), nb, gs, 0, s, n, deg, off, col, o16, c16, val, x, y, rm); break;
      case 5: hipLaunchKernelGGL((k_spmv<5, T, X>), nb, gs, 0, s, n, deg, off, col, o16, c16, val, x, y, rm); break;
      case 6: hipLaunchKernelGGL((k_spmv<6, T, X>), nb, gs, 0, s, n, deg, off, col, o16, c16, val, x, y, rm); break;
      case 7: hipLaunchKernelGGL((k_spmv<7, T, X>), nb, gs, 0, s, n, deg, off, col, o16, c16, val, x, y, rm); break;
      case 8: hipLaunchKernelGGL((k_spmv<8, T, X>), nb, gs, 0, s, n, deg, off, col, o16, c16, val, x, y, rm); break;
      default: return -1;
   }
   return hipGetLastError() == hipSuccess ? 1 : -6;
}

static Vector *Vec(GmlCtx *g, int i) { return (i >= 1 && i <= GmlMaxVec) ? g->vec[i] : nullptr; }
static Matrix *Mat(GmlCtx *g, int i) { return (i >= 1 && i <= GmlMaxMat) ? g->mat[i] : nullptr; }

// ===========================================================================
// Block-Jacobi preconditioned CG (FP32 vectors, FP64 dot accumulation), and
// iterative refinement with the residual b - A x evaluated in FP64.
// One iteration is three passes:
//   k_cg_spmv:    q = A p, p.q -> alpha = rz / pq
//   k_cg_update:  x += alpha p, r -= alpha q, z = Minv r, r.z, r.r -> beta
//   k_cg_pupdate: p = z + beta p
// (Forming p = z + beta p inside the SpMV, where it is gathered, measured
// slower: it doubles the gathered loads of the product.)
// Grid sums are deterministic and fence-free: each work-group stores its
// partial with a plain store, and a one-work-group kernel (k_cg_finish) adds
// the partials in block order right after; the kernel boundary orders the
// two.  (A last-block ticket would need a device-scope fence per work-group,
// which on CDNA writes back and invalidates the XCD's L2 every time.)
// All scalars stay on the device; the host polls r.r every few iterations
// without stalling the stream.
// ===========================================================================
static const int kCgBlock = 256;
static const int kCgGrid = 4096;

struct CgScal {
   double rz, pq, rr, alpha, beta, rz0, rr0;
};

// per-block partial sums: 2 slots of np doubles each
struct CgPart {
   double *p;
   unsigned int np;
};

__device__ __forceinline__ double wsum(double v)
{
   for(int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
   return v;
}

// this work-group's sums of a (into slot 0) and b (into slot 1 if two)
__device__ __forceinline__ void block_sums(double a, double b, CgPart P, bool two)
{
   __shared__ double lds[2][kCgBlock / 64];
   a = wsum(a);
   b = wsum(b);
   if((threadIdx.x & 63) == 0) { lds[0][threadIdx.x >> 6] = a; lds[1][threadIdx.x >> 6] = b; }
   __syncthreads();
   if(threadIdx.x == 0)
   {
      double s = 0, t = 0;
      for(int i = 0; i < (int)(blockDim.x >> 6); i++) { s += lds[0][i]; t += lds[1][i]; }
      P.p[blockIdx.x] = s;
      if(two) P.p[P.np + blockIdx.x] = t;
   }
}

// sum of n partials in a fixed order, by one work-group of kFinBlock threads,
// eight independent loads in flight per thread
static const int kFinBlock = 1024;

__device__ double fin_sum(const double *part, unsigned int n)
{
   __shared__ double lds[kFinBlock / 64];
   double s = 0;
   for(unsigned int base = 0; base < n; base += 8 * blockDim.x)
   {
      double v[8];
#pragma unroll
      for(int u = 0; u < 8; u++)
      {
         const unsigned int i = base + u * blockDim.x + threadIdx.x;
         v[u] = i < n ? part[i] : 0.;
      }
#pragma unroll
      for(int u = 0; u < 8; u++) s += v[u];
   }
   s = wsum(s);
   __syncthreads();
   if((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
   __syncthreads();
   s = 0;
   for(int i = 0; i < (int)(blockDim.x >> 6); i++) s += lds[i];
   return s;
}

enum { CgFinAlpha, CgFinBeta, CgFinInit, CgFinResid, CgFinDot };

// one work-group: finish the grid sums of the kernel just run (n partials)
template<int W>
__global__ void __launch_bounds__(kFinBlock) k_cg_finish(CgScal *S, CgPart P, unsigned int n)
{
   const double s0 = fin_sum(P.p, n);
   const double s1 = (W == CgFinBeta || W == CgFinInit) ? fin_sum(P.p + P.np, n) : 0.;
   if(threadIdx.x) return;
   if(W == CgFinAlpha) { S->pq = s0; S->alpha = s0 == 0. ? 0. : S->rz / s0; }
   if(W == CgFinBeta)  { S->beta = S->rz == 0. ? 0. : s0 / S->rz; S->rz = s0; S->rr = s1; }
   if(W == CgFinInit)  { S->rz = s0; S->rz0 = s0; S->rr = s1; S->rr0 = s1; S->beta = 0.; }
   if(W == CgFinResid) S->rr0 = s0;
   if(W == CgFinDot)   S->pq = s0;
}

// q = A p and this work-group's part of p.q
template<int B>
__global__ void __launch_bounds__(256) k_cg_spmv(int n, const int *__restrict__ deg, const unsigned int *__restrict__ off,
                                                 const int *__restrict__ col, const unsigned int *__restrict__ off16,
                                                 const short *__restrict__ col16, const float *__restrict__ val,
                                                 const float *__restrict__ p, float *__restrict__ q, CgPart P, int remap)
{
   const int r = (int)(xcd_block(remap) * blockDim.x + threadIdx.x);
   double pq = 0;
   if(r < n)
   {
      float acc[B];
#pragma unroll
      for(int i = 0; i < B; i++) acc[i] = 0.f;
      const unsigned int b16 = off16[r >> 6];
      row_product<B, float, float>(r, r & 63, deg[r], (size_t)off[r >> 6], (int)(off16[(r >> 6) + 1] - b16), (size_t)b16,
                                   col, col16, val, p, acc);
#pragma unroll
      for(int i = 0; i < B; i++)
      {
         const size_t e = (size_t)r * B + i;
         q[e] = acc[i];
         pq += (double)p[e] * (double)acc[i];
      }
   }
   block_sums(pq, 0., P, false);
}

// p = z + beta p, four entries per thread
__global__ void __launch_bounds__(256) k_cg_pupdate(long long n, float *__restrict__ p, const float *__restrict__ z,
                                                    const CgScal *S)
{
   const float b = (float)S->beta;
   const long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * 4;
   if(i + 3 < n)
   {
      float4 pv = *(const float4 *)&p[i], zv = *(const float4 *)&z[i];
      pv.x = zv.x + b * pv.x; pv.y = zv.y + b * pv.y; pv.z = zv.z + b * pv.z; pv.w = zv.w + b * pv.w;
      *(float4 *)&p[i] = pv;
   }
   else
      for(long long k = i; k < n; k++) p[k] = z[k] + b * p[k];
}

// x += alpha p; r -= alpha q; z = Minv r (Minv entry e of row i at e*nrow+i); r.z, r.r -> beta
template<int B>
__global__ void __launch_bounds__(kCgBlock) k_cg_update(int nrow, float *__restrict__ x, float *__restrict__ r,
                                                        const float *__restrict__ p, const float *__restrict__ q,
                                                        const float *__restrict__ Minv, float *__restrict__ z,
                                                        const CgScal *S, CgPart P)
{
   const float a = (float)S->alpha;
   double rz = 0, rr = 0;
   for(int row = blockIdx.x * blockDim.x + threadIdx.x; row < nrow; row += gridDim.x * blockDim.x)
   {
      float rv[B];
#pragma unroll
      for(int i = 0; i < B; i++)
      {
         size_t k = (size_t)row * B + i;
         x[k] += a * p[k];
         rv[i] = r[k] - a * q[k];
         r[k] = rv[i];
      }
#pragma unroll
      for(int i = 0; i < B; i++)
      {
         float t = 0.f;
#pragma unroll
         for(int m = 0; m < B; m++) t += Minv[(size_t)(i * B + m) * nrow + row] * rv[m];
         z[(size_t)row * B + i] = t;
         rz += (double)rv[i] * (double)t;
         rr += (double)rv[i] * (double)rv[i];
      }
   }
   block_sums(rz, rr, P, true);
}

// init: r = rhs (x = 0), z = Minv r, p = z, beta = 0, rz = r.z, rr = r.r
template<int B>
__global__ void __launch_bounds__(kCgBlock) k_cg_init(int nrow, const float *__restrict__ rhs, float *__restrict__ x, float *__restrict__ r,
                                                      float *__restrict__ p, float *__restrict__ z,
                                                      const float *__restrict__ Minv, CgPart P)
{
   double rz = 0, rr = 0;
   for(int row = blockIdx.x * blockDim.x + threadIdx.x; row < nrow; row += gridDim.x * blockDim.x)
   {
      float rv[B];
#pragma unroll
      for(int i = 0; i < B; i++)
      {
         size_t k = (size_t)row * B + i;
         rv[i] = rhs[k];
         r[k] = rv[i];
         x[k] = 0.f;
      }
#pragma unroll
      for(int i = 0; i < B; i++)
      {
         float t = 0.f;
#pragma unroll
         for(int m = 0; m < B; m++) t += Minv[(size_t)(i * B + m) * nrow + row] * rv[m];
         z[(size_t)row * B + i] = t;
         p[(size_t)row * B + i] = t;
         rz += (double)rv[i] * (double)t;
         rr += (double)rv[i] * (double)rv[i];
      }
   }
   block_sums(rz, rr, P, true);
}

// block-Jacobi: invert the diagonal block of every row (FP64 Gauss-Jordan
// with partial pivoting; every loop unrolled and the row swap done by
// selects, so the B x 2B tableau stays in registers); entry e of row r's
// inverse lands at Minv[e * nrow + r]
template<int B>
__global__ void __launch_bounds__(256) k_block_jacobi(int nrow, const int *__restrict__ deg, const unsigned int *__restrict__ off,
                                                      const int *__restrict__ col, const float *__restrict__ val,
                                                      float *__restrict__ Minv)
{
   int r = blockIdx.x * blockDim.x + threadIdx.x;
   if(r >= nrow) return;
   const int l = r & 63, d = deg[r];
   const size_t base = off[r >> 6];
   double a[B][2 * B];
#pragma unroll
   for(int i = 0; i < B; i++)
#pragma unroll
      for(int j = 0; j < 2 * B; j++) a[i][j] = (j - B == i) ? 1. : 0.;
   for(int k = 0; k < d; k++)
      if(col[(base + k) * 64 + l] == r)
      {
#pragma unroll
         for(int i = 0; i < B; i++)
#pragma unroll
            for(int j = 0; j < B; j++) a[i][j] += val[((base + k) * B * B + i * B + j) * 64 + l];
      }
#pragma unroll
   for(int c = 0; c < B; c++)
   {
      // pivot: the largest |a[i][c]|, i >= c, first one on ties
      double best = fabs(a[c][c]);
      int piv = c;
#pragma unroll
      for(int i = c + 1; i < B; i++)
         if(fabs(a[i][c]) > best) { best = fabs(a[i][c]); piv = i; }
#pragma unroll
      for(int i = c + 1; i < B; i++)
      {
         const bool sw = (i == piv);
#pragma unroll
         for(int j = 0; j < 2 * B; j++)
         {
            const double t = a[c][j];
            a[c][j] = sw ? a[i][j] : t;
            a[i][j] = sw ? t : a[i][j];
         }
      }
      const double inv = a[c][c] != 0. ? 1. / a[c][c] : 0.;
#pragma unroll
      for(int j = 0; j < 2 * B; j++) a[c][j] *= inv;
#pragma unroll
      for(int i = 0; i < B; i++)
         if(i != c)
         {
            const double f = a[i][c];
#pragma unroll
            for(int j = 0; j < 2 * B; j++) a[i][j] -= f * a[c][j];
         }
   }
#pragma unroll
   for(int i = 0; i < B; i++)
#pragma unroll
      for(int j = 0; j < B; j++) Minv[(size_t)(i * B + j) * nrow + r] = (float)a[i][B + j];
}

// FP64 residual r64 = b - A x64 (grid-stride over rows) and its squared norm
template<int B>
__global__ void __launch_bounds__(kCgBlock) k_resid64(int nrow, const int *__restrict__ deg, const unsigned int *__restrict__ off,
                                                      const int *__restrict__ col, const float *__restrict__ val,
                                                      const double *__restrict__ x, const float *__restrict__ rhs,
                                                      double *__restrict__ r64, CgPart P)
{
   double nrm = 0;
   for(int r = blockIdx.x * blockDim.x + threadIdx.x; r < nrow; r += gridDim.x * blockDim.x)
   {
      const int l = r & 63, d = deg[r];
      const size_t base = off[r >> 6];
      double acc[B];
      for(int i = 0; i < B; i++) acc[i] = rhs[(size_t)r * B + i];
      for(int k = 0; k < d; k++)
      {
         const size_t s = base + k;
         const int j = col[s * 64 + l];
         for(int i = 0; i < B; i++)
            for(int m = 0; m < B; m++)
               acc[i] -= (double)val[(s * B * B + i * B + m) * 64 + l] * x[(size_t)j * B + m];
      }
      for(int i = 0; i < B; i++)
      {
         r64[(size_t)r * B + i] = acc[i];
         nrm += acc[i] * acc[i];
      }
   }
   block_sums(nrm, 0., P, false);
}

__global__ void k_axpy64(long long n, double *__restrict__ x, const float *__restrict__ d, double a)
{
   long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(i < n) x[i] += a * (double)d[i];
}

__global__ void k_f64_to_f32(long long n, const double *__restrict__ a, float *__restrict__ b)
{
   long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(i < n) b[i] = (float)a[i];
}

__global__ void __launch_bounds__(kCgBlock) k_dot(long long n, const float *__restrict__ a, const float *__restrict__ b,
                                                  CgPart P)
{
   double acc = 0;
   for(long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
      acc += (double)a[i] * (double)b[i];
   block_sums(acc, 0., P, false);
}

// CG workspace of a matrix: allocated by its first solve and kept (no
// allocation or free, hence no device-wide synchronisation, in later
// solves); the block-Jacobi inverse is computed once
struct CgWork {
   float *x = nullptr, *r = nullptr, *p = nullptr, *q = nullptr, *z = nullptr, *Minv = nullptr, *d32 = nullptr;
   double *x64 = nullptr, *r64 = nullptr;
   CgScal *S = nullptr;
   CgPart P = {nullptr, 0};
   double *host = nullptr;   // pinned: rr readback
   bool minv_ok = false;
   size_t bytes = 0;
};

static void FreeCgWork(GmlCtx *g, CgWork *w)
{
   if(!w) return;
   DevFree(g, w->x); DevFree(g, w->r); DevFree(g, w->p); DevFree(g, w->q); DevFree(g, w->z);
   DevFree(g, w->d32); DevFree(g, w->Minv); DevFree(g, w->x64); DevFree(g, w->r64); DevFree(g, w->S);
   DevFree(g, w->P.p);
   if(w->host) (void)hipHostFree(w->host);
   g->MemSiz -= w->bytes;
   delete w;
}

static CgWork *CgWorkspace(GmlCtx *g, Matrix *A)
{
   if(A->cg) return A->cg;
   const int nrow = A->NmbLin, B = A->b;
   const long long n = (long long)nrow * B;
   CgWork *w = new CgWork();
   w->P.np = std::max<unsigned int>(kCgGrid, Blocks(nrow, 64));
   bool ok = true;
   auto al = [&](size_t by) { void *p = DevAlloc(g, by); if(!p) ok = false; w->bytes += by; return p; };
   w->x = (float *)al(n * 4); w->r = (float *)al(n * 4); w->p = (float *)al(n * 4);
   w->q = (float *)al(n * 4); w->z = (float *)al(n * 4); w->d32 = (float *)al(n * 4);
   w->Minv = (float *)al((size_t)nrow * B * B * 4);
   w->x64 = (double *)al(n * 8); w->r64 = (double *)al(n * 8);
   w->S = (CgScal *)al(sizeof(CgScal));
   w->P.p = (double *)al((size_t)2 * w->P.np * 8);
   if(ok && hipHostMalloc((void **)&w->host, 64, hipHostMallocDefault) != hipSuccess) ok = false;
   g->MemSiz += w->bytes;
   if(!ok) { FreeCgWork(g, w); return nullptr; }
   A->cg = w;
   return w;
}

} // namespace gml

using namespace gml;

extern "C" {

int GmlNewMatrix(size_t GmlIdx, int NmbLin, int NmbBlk, int BlkSiz, void *val, int *col, int *lin, int FltTyp)
{
   GML_CTX(g, GmlIdx);
   if(NmbLin < 1 || NmbBlk < 1 || BlkSiz < 1 || BlkSiz > 8 || !val || !col || !lin)
      return 0;
   if(FltTyp != GmlFlt && FltTyp != GmlDbl) return 0;
   if(lin[0] != 0 || lin[NmbLin] != NmbBlk) return 0;
   int idx = 0;
   for(int i = 1; i <= GmlMaxMat; i++) if(!g->mat[i]) { idx = i; break; }
   if(!idx) return 0;
   const int b = BlkSiz, bb = b * b, fs = FltTyp == GmlFlt ? 4 : 8;
   const long long nchunk = (NmbLin + 63) / 64;
   std::vector<unsigned int> off((size_t)nchunk + 1, 0);
   std::vector<int> deg(NmbLin);
   for(int r = 0; r < NmbLin; r++)
   {
      deg[r] = lin[r + 1] - lin[r];
      if(deg[r] < 0) return 0;
      for(int k = lin[r]; k < lin[r + 1]; k++)
         if(col[k] < 0 || col[k] >= NmbLin) return 0;
   }
   for(long long c = 0; c < nchunk; c++)
   {
      int w = 0;
      for(long long r = c * 64; r < std::min<long long>(NmbLin, c * 64 + 64); r++) w = std::max(w, deg[r]);
      off[c + 1] = off[c] + w;
   }
   long long nslab = off[nchunk];
   // each row's blocks in slot order: those whose column lies within +-32767
   // of the row first, then the others, each group in the caller's order;
   // a slab's first kn slots (kn = its rows' smallest count of near blocks)
   // also get their columns as 16-bit offsets from the row
   auto near = [](int c, int r) { return c - r <= 32767 && r - c <= 32767; };
   std::vector<unsigned int> off16((size_t)nchunk + 1, 0);
   for(long long c = 0; c < nchunk; c++)
   {
      int kn = INT32_MAX;
      for(long long r = c * 64; r < std::min<long long>(NmbLin, c * 64 + 64); r++)
      {
         int nn = 0;
         for(int k = lin[r]; k < lin[r + 1]; k++) nn += near(col[k], (int)r);
         kn = std::min(kn, nn);
      }
      off16[c + 1] = off16[c] + (unsigned int)kn;
   }
   const long long nslab16 = off16[nchunk];
   std::vector<int> hc((size_t)nslab * 64 + 64, 0);
   std::vector<short> h16((size_t)nslab16 * 64 + 64, 0);
   std::vector<char> hv(((size_t)nslab * bb * 64 + 64) * fs, 0);
   std::vector<int> order;
   for(int r = 0; r < NmbLin; r++)
   {
      const size_t base = off[r >> 6], base16 = off16[r >> 6], l = r & 63;
      const int kn = (int)(off16[(r >> 6) + 1] - off16[r >> 6]);
      order.clear();
      for(int k = lin[r]; k < lin[r + 1]; k++) if(near(col[k], r)) order.push_back(k);
      for(int k = lin[r]; k < lin[r + 1]; k++) if(!near(col[k], r)) order.push_back(k);
      for(int k = 0; k < deg[r]; k++)
      {
         const size_t s = base + k;
         const int src = order[k];
         hc[s * 64 + l] = col[src];
         if(k < kn) h16[(base16 + k) * 64 + l] = (short)(col[src] - r);
         for(int e = 0; e < bb; e++)
            memcpy(&hv[((s * bb + e) * 64 + l) * fs], (char *)val + ((size_t)src * bb + e) * fs, fs);
      }
   }
   Matrix *m = new Matrix();
   m->NmbLin = NmbLin; m->NmbBlk = NmbBlk; m->b = b; m->flt = FltTyp; m->nslab = nslab;
   m->nslab16 = nslab16;
   for(int r = 0; r < NmbLin; r++) m->n16 += off16[(r >> 6) + 1] - off16[r >> 6];
   m->deg = (int *)DevAlloc(g, (size_t)NmbLin * 4);
   m->off = (unsigned int *)DevAlloc(g, off.size() * 4);
   m->col = (int *)DevAlloc(g, hc.size() * 4);
   m->val = DevAlloc(g, hv.size());
   m->off16 = (unsigned int *)DevAlloc(g, off16.size() * 4);
   m->col16 = (short *)DevAlloc(g, h16.size() * 2);
   if(!m->deg || !m->off || !m->col || !m->val || !m->off16 || !m->col16
   || !DevCopy(g, m->deg, deg.data(), (size_t)NmbLin * 4, 1)
   || !DevCopy(g, m->off, off.data(), off.size() * 4, 1)
   || !DevCopy(g, m->col, hc.data(), hc.size() * 4, 1)
   || !DevCopy(g, m->val, hv.data(), hv.size(), 1)
   || !DevCopy(g, m->off16, off16.data(), off16.size() * 4, 1)
   || !DevCopy(g, m->col16, h16.data(), h16.size() * 2, 1))
   {
      DevFree(g, m->deg); DevFree(g, m->off); DevFree(g, m->col); DevFree(g, m->val);
      DevFree(g, m->off16); DevFree(g, m->col16);
      delete m;
      return 0;
   }
   const size_t held = (size_t)NmbLin * 4 + off.size() * 4 + hc.size() * 4 + hv.size() + off16.size() * 4 + h16.size() * 2;
   g->MemSiz += held;
   g->MovSiz += held;
   // bytes and flops of one product, as the reference models them
   m->MemAcc = ((double)NmbLin * b + (double)NmbBlk * b * (b + 1)) * fs;
   m->FltOpp = (double)NmbBlk * b * b * 2;
   g->mat[idx] = m;
   DbgPrint(g, "Matrix %d: %d block rows, %d blocks of %dx%d, %lld SELL-64 slabs (%.1f%% occupancy), "
            "%lld with 16-bit columns\n", idx, NmbLin, NmbBlk, b, b, nslab, nslab ? 100. * NmbBlk / (nslab * 64.) : 0.,
            nslab16);
   return idx;
}

int GmlNewVector(size_t GmlIdx, int NmbLin, int BlkSiz, void *val, int FltTyp)
{
   GML_CTX(g, GmlIdx);
   if(NmbLin < 1 || BlkSiz < 1 || BlkSiz > 64) return 0;
   if(FltTyp != GmlFlt && FltTyp != GmlDbl) return 0;
   int idx = 0;
   for(int i = 1; i <= GmlMaxVec; i++) if(!g->vec[i]) { idx = i; break; }
   if(!idx) return 0;
   Vector *v = new Vector();
   v->NmbLin = NmbLin; v->b = BlkSiz; v->flt = FltTyp;
   v->bytes = (size_t)NmbLin * BlkSiz * (FltTyp == GmlFlt ? 4 : 8);
   v->dev = DevAlloc(g, v->bytes);
   if(!v->dev || (val && !DevCopy(g, v->dev, val, v->bytes, 1)))
   {
      DevFree(g, v->dev);
      delete v;
      return 0;
   }
   if(!val && !g->host_only) HIPCHK(hipMemsetAsync(v->dev, 0, v->bytes, g->stream));
   g->MemSiz += v->bytes;
   if(val) g->MovSiz += v->bytes;
   g->vec[idx] = v;
   return idx;
}

int GmlMultMatVec(size_t GmlIdx, int MatIdx, int VecIdx1, int VecIdx2)
{
   GML_CTX(g, GmlIdx);
   Matrix *m = Mat(g, MatIdx);
   if(!m) { printf("Invalid data index: %d\n", MatIdx); return -1; }
   Vector *y = Vec(g, VecIdx1), *x = Vec(g, VecIdx2);
   if(!y) { printf("Invalid data index: %d\n", VecIdx1); return -2; }
   if(!x) { printf("Invalid data index: %d\n", VecIdx2); return -3; }
   if(m->NmbLin != y->NmbLin || m->NmbLin != x->NmbLin || y->b != m->b || x->b != m->b || x->flt != y->flt)
   {
      printf("vectors and matrix size differ: %d(%d) %d(%d) %d(%d)\n",
             MatIdx, m->NmbLin, VecIdx1, y->NmbLin, VecIdx2, x->NmbLin);
      return -4;
   }
   if(g->host_only) return -6;
   if(m->flt == GmlDbl && x->flt != GmlDbl) return -4;
   int rc = 1;
   int tl = TimedLaunch(g, GmlxBuiltinSpMV, [&]() {
      if(m->flt == GmlFlt && x->flt == GmlFlt) rc = SpMV<float, float>(g, m, (const float *)x->dev, (float *)y->dev);
      else if(m->flt == GmlFlt)                rc = SpMV<float, double>(g, m, (const double *)x->dev, (double *)y->dev);
      else                                     rc = SpMV<double, double>(g, m, (const double *)x->dev, (double *)y->dev);
   });
   if(tl != 1) return tl;
   if(rc != 1) return rc;
   g->MemAcc += m->MemAcc;
   g->FltOpp += m->FltOpp;
   return 1;
}

int GmlMultDiagMatVec(size_t GmlIdx, int VecIdx1, int VecIdx2, int VecIdx3)
{
   GML_CTX(g, GmlIdx);
   Vector *D = Vec(g, VecIdx1), *u = Vec(g, VecIdx2), *v = Vec(g, VecIdx3);
   if(!D) { printf("Invalid data index: %d\n", VecIdx1); return -1; }
   if(!u) { printf("Invalid data index: %d\n", VecIdx2); return -2; }
   if(!v) { printf("Invalid data index: %d\n", VecIdx3); return -3; }
   if(D->NmbLin != v->NmbLin || D->NmbLin != u->NmbLin || D->b != v->b * v->b || u->b != v->b
   || D->flt != u->flt || u->flt != v->flt)
   {
      printf("vector sizes differ: ID %d (%d x %d) and ID %d (%d x %d)\n",
             VecIdx1, D->NmbLin, D->b, VecIdx3, v->NmbLin, v->b);
      return -4;
   }
   if(g->host_only) return -6;
   unsigned int nb = Blocks(D->NmbLin);
   if(D->flt == GmlFlt)
      hipLaunchKernelGGL(k_diagmul<float>, dim3(nb), dim3(256), 0, g->stream, D->NmbLin, v->b,
                         (const float *)D->dev, (const float *)u->dev, (float *)v->dev);
   else
      hipLaunchKernelGGL(k_diagmul<double>, dim3(nb), dim3(256), 0, g->stream, D->NmbLin, v->b,
                         (const double *)D->dev, (const double *)u->dev, (double *)v->dev);
   if(hipGetLastError() != hipSuccess) return -6;
   int fs = D->flt == GmlFlt ? 4 : 8;
   g->MemAcc += (double)D->NmbLin * (D->b + 2 * u->b) * fs;
   g->FltOpp += (double)D->NmbLin * 2 * D->b;
   return 1;
}

int GmlAddVec3(size_t GmlIdx, int VecIdx1, int VecIdx2, int VecIdx3, int VecIdx4)
{
   GML_CTX(g, GmlIdx);
   Vector *a = Vec(g, VecIdx1), *b = Vec(g, VecIdx2), *c = Vec(g, VecIdx3), *y = Vec(g, VecIdx4);
   if(!a) { printf("Invalid data index: %d\n", VecIdx1); return -1; }
   if(!b) { printf("Invalid data index: %d\n", VecIdx2); return -2; }
   if(!c) { printf("Invalid data index: %d\n", VecIdx3); return -3; }
   if(!y) { printf("Invalid data index: %d\n", VecIdx4); return -4; }
   if(a->NmbLin != b->NmbLin || a->b != b->b || a->NmbLin != c->NmbLin || a->b != c->b || a->flt != b->flt || a->flt != c->flt)
      return -4;
   if(a->NmbLin != y->NmbLin || a->b != y->b || a->flt != y->flt)
      return -5;
   if(g->host_only) return -6;
   long long n = (long long)a->NmbLin * a->b;
   if(a->flt == GmlFlt)
      hipLaunchKernelGGL(k_add3<float>, dim3(VecBlocks<float>(n)), dim3(256), 0, g->stream, n, (const float *)a->dev,
                         (const float *)b->dev, (const float *)c->dev, (float *)y->dev);
   else
      hipLaunchKernelGGL(k_add3<double>, dim3(VecBlocks<double>(n)), dim3(256), 0, g->stream, n, (const double *)a->dev,
                         (const double *)b->dev, (const double *)c->dev, (double *)y->dev);
   if(hipGetLastError() != hipSuccess) return -6;
   g->MemAcc += (double)a->NmbLin * a->b * (a->flt == GmlFlt ? 4 : 8) * 3;
   g->FltOpp += (double)a->NmbLin * a->b;
   return 1;
}

// The reference takes the factor from the user parameter structure
// (GmlPar->scale, a float after an int and a float) and ignores mul; here an
// explicit *mul wins and the parameter field is used when mul is NULL.
int GmlScaleVec(size_t GmlIdx, int VecIdx, double *mul)
{
   GML_CTX(g, GmlIdx);
   Vector *v = Vec(g, VecIdx);
   if(!v) { printf("Invalid data index: %d\n", VecIdx); return -2; }
   double s;
   if(mul) s = *mul;
   else if(g->ParIdx && g->dat[g->ParIdx].MemSiz >= 12 && g->dat[g->ParIdx].host)
      s = ((const float *)g->dat[g->ParIdx].host)[2];
   else return -3;
   if(g->host_only) return -6;
   long long n = (long long)v->NmbLin * v->b;
   if(v->flt == GmlFlt)
      hipLaunchKernelGGL(k_scale<float>, dim3(VecBlocks<float>(n)), dim3(256), 0, g->stream, n, (float *)v->dev, (float)s);
   else
      hipLaunchKernelGGL(k_scale<double>, dim3(VecBlocks<double>(n)), dim3(256), 0, g->stream, n, (double *)v->dev, s);
   if(hipGetLastError() != hipSuccess) return -6;
   g->MemAcc += (double)v->NmbLin * v->b * (v->flt == GmlFlt ? 4 : 8) * 2;
   g->FltOpp += (double)v->NmbLin * v->b;
   return 1;
}

// sqrt of the sum over rows of each row's 2-norm (the reference's meaning):
// the row norms land in the float data RedIdx, then GmlReduceVector(L1).
int GmlNormVec(size_t GmlIdx, int VecIdx, int RedIdx, double *nrm)
{
   GML_CTX(g, GmlIdx);
   Vector *v = Vec(g, VecIdx);
   if(!v) { printf("Invalid data index: %d\n", VecIdx); return -2; }
   if(RedIdx < 1 || RedIdx > GmlMaxDat || !g->dat[RedIdx].use || g->dat[RedIdx].ItmTyp != GmlFlt
   || g->dat[RedIdx].NmbItm != 1 || g->dat[RedIdx].NmbLin < v->NmbLin)
      return -3;
   if(g->host_only) return -6;
   float *out = (float *)g->dat[RedIdx].dev;
   if(v->flt == GmlFlt)
      hipLaunchKernelGGL(k_rownorm<float>, dim3(Blocks(v->NmbLin)), dim3(256), 0, g->stream, v->NmbLin, v->b,
                         (const float *)v->dev, out);
   else
      hipLaunchKernelGGL(k_rownorm<double>, dim3(Blocks(v->NmbLin)), dim3(256), 0, g->stream, v->NmbLin, v->b,
                         (const double *)v->dev, out);
   if(hipGetLastError() != hipSuccess) return -6;
   g->dat[RedIdx].host_valid = false;
   BumpVersion(&g->dat[RedIdx]);
   double s = 0.;
   int rc = ReduceDevice(g, out, v->NmbLin, GmlL1, &s, true);
   if(rc != 1) return rc;
   *nrm = sqrt(s);
   g->MemAcc += (double)v->NmbLin * v->b * (v->flt == GmlFlt ? 4 : 8) * 2;
   g->FltOpp += (double)v->NmbLin * v->b * 2;
   return 1;
}

int GmlxSetVector(size_t GmlIdx, int VecIdx, const void *val)
{
   GML_CTX(g, GmlIdx);
   Vector *v = Vec(g, VecIdx);
   if(!v || !val) return 0;
   return DevCopy(g, v->dev, val, v->bytes, 1) ? 1 : 0;
}

int GmlxGetVector(size_t GmlIdx, int VecIdx, void *val)
{
   GML_CTX(g, GmlIdx);
   Vector *v = Vec(g, VecIdx);
   if(!v || !val) return 0;
   return DevCopy(g, val, v->dev, v->bytes, 2) ? 1 : 0;
}

int GmlxFreeVector(size_t GmlIdx, int VecIdx)
{
   GML_CTX(g, GmlIdx);
   Vector *v = Vec(g, VecIdx);
   if(!v) return 0;
   DevFree(g, v->dev);
   g->MemSiz -= v->bytes;
   delete v;
   g->vec[VecIdx] = nullptr;
   return 1;
}

int GmlxGetMatrixInfo(size_t GmlIdx, int MatIdx, long long *info)
{
   GML_CTX(g, GmlIdx);
   const Matrix *m = Mat(g, MatIdx);
   if(!m || !info) return 0;
   info[0] = m->NmbLin; info[1] = m->NmbBlk; info[2] = m->b;
   info[3] = m->nslab; info[4] = m->nslab16; info[5] = m->n16;
   return 1;
}

int GmlxFreeMatrix(size_t GmlIdx, int MatIdx)
{
   GML_CTX(g, GmlIdx);
   Matrix *m = Mat(g, MatIdx);
   if(!m) return 0;
   FreeCgWork(g, m->cg);
   DevFree(g, m->deg); DevFree(g, m->col); DevFree(g, m->off); DevFree(g, m->val);
   DevFree(g, m->off16); DevFree(g, m->col16);
   delete m;
   g->mat[MatIdx] = nullptr;
   return 1;
}

// Block CG on A x = rhs (FP32 matrix and right-hand side; the solution
// vector may be FP32 or FP64).  refine = 0: one FP32
// Jacobi-PCG run to tol (relative to |rhs|, on the recurrence residual).
// refine = 1: iterative refinement -- x kept in FP64, the residual rhs - A x
// evaluated with FP64 accumulation, each correction solved in FP32 to a
// relative 1e-4 (or inner_tol) -- until the FP64 residual meets tol.  On
// return *relres is the FP64 relative residual |rhs - A x| / |rhs| and
// *iters the total inner iterations.  Returns 1, or a negative error.
int GmlxSolveCG(size_t GmlIdx, int MatIdx, int RhsIdx, int SolIdx, double tol, int maxit, int refine,
                double *relres, int *iters)
{
   GML_CTX(g, GmlIdx);
   Matrix *A = Mat(g, MatIdx);
   Vector *b = Vec(g, RhsIdx), *xo = Vec(g, SolIdx);
   if(!A || !b || !xo) return -1;
   if(A->flt != GmlFlt || b->flt != GmlFlt || b->b != A->b || xo->b != A->b
   || b->NmbLin != A->NmbLin || xo->NmbLin != A->NmbLin)
      return -4;
   if(A->b != 3 && A->b != 1 && A->b != 4) return -4;
   if(g->host_only) return -6;
   const int nrow = A->NmbLin, B = A->b;
   const long long n = (long long)nrow * B;
   hipStream_t s = g->stream;
   const Shape sh = BuiltinShape(g, GmlxBuiltinCgSpMV), su = BuiltinShape(g, GmlxBuiltinCgUpdate);
   const unsigned int nbs = Blocks(nrow, sh.grp);
   CgWork *ws = CgWorkspace(g, A);
   if(!ws) return -5;
   CgWork &w = *ws;
   auto cleanup = []() {};
   HIPCHK(hipMemsetAsync(w.S, 0, sizeof(CgScal), s));
   HIPCHK(hipMemsetAsync(w.x64, 0, n * 8, s));
   unsigned int nbr = Blocks(nrow), nbn = Blocks(n);
   unsigned int gr = std::min<unsigned int>(kCgGrid, Blocks(nrow, kCgBlock));
   unsigned int gu = std::min<unsigned int>(kCgGrid, Blocks(nrow, su.grp));
   unsigned int gn = std::min<unsigned int>(kCgGrid, Blocks(n, kCgBlock));
#define GML_B(K, ...) do { switch(B) { case 1: hipLaunchKernelGGL(K<1>, __VA_ARGS__); break; \
   case 3: hipLaunchKernelGGL(K<3>, __VA_ARGS__); break; default: hipLaunchKernelGGL(K<4>, __VA_ARGS__); } } while(0)
   if(!w.minv_ok)
   {
      GML_B(k_block_jacobi, dim3(nbr), dim3(256), 0, s, nrow, A->deg, A->off, A->col, (const float *)A->val, w.Minv);
      w.minv_ok = true;
   }

   // |rhs|^2
   hipLaunchKernelGGL(k_dot, dim3(gn), dim3(kCgBlock), 0, s, n, (const float *)b->dev, (const float *)b->dev, w.P);
   hipLaunchKernelGGL(k_cg_finish<CgFinDot>, dim3(1), dim3(kFinBlock), 0, s, w.S, w.P, gn);
   HIPCHK(hipMemcpyAsync(w.host, &w.S->pq, 8, hipMemcpyDeviceToHost, s));
   HIPCHK(hipStreamSynchronize(s));
   const double bnorm = sqrt(w.host[0]);
   if(bnorm == 0.)
   {
      HIPCHK(hipMemsetAsync(xo->dev, 0, n * (xo->flt == GmlDbl ? 8 : 4), s));
      HIPCHK(hipStreamSynchronize(s));
      *relres = 0.; *iters = 0;
      cleanup();
      return 1;
   }

   // one inner PCG solve of A d = rhs32 to relative tol itol, d in w.x
   int total = 0;
   auto pcg = [&](const float *rhs32, double itol, int cap) -> int {
      GML_B(k_cg_init, dim3(gr), dim3(kCgBlock), 0, s, nrow, rhs32, w.x, w.r, w.p, w.z, w.Minv, w.P);
      hipLaunchKernelGGL(k_cg_finish<CgFinInit>, dim3(1), dim3(kFinBlock), 0, s, w.S, w.P, gr);
      HIPCHK(hipMemcpyAsync(w.host, &w.S->rr0, 8, hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      const double r0 = w.host[0];
      const double goal = itol * itol * r0;
      hipEvent_t ev;
      HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      int it = 0, pending = -1;
      double rr = r0;
      while(it < cap && rr > goal)
      {
         const int chunk = 8;
         for(int c = 0; c < chunk && it < cap; c++, it++)
         {
            int t1 = TimedLaunch(g, GmlxBuiltinCgSpMV, [&]() {
               GML_B(k_cg_spmv, dim3(nbs), dim3(sh.grp), 0, s, nrow, A->deg, A->off, A->col, A->off16, A->col16,
                     (const float *)A->val, w.p, w.q, w.P, sh.remap);
               hipLaunchKernelGGL(k_cg_finish<CgFinAlpha>, dim3(1), dim3(kFinBlock), 0, s, w.S, w.P, nbs);
            });
            int t2 = TimedLaunch(g, GmlxBuiltinCgUpdate, [&]() {
               GML_B(k_cg_update, dim3(gu), dim3(su.grp), 0, s, nrow, w.x, w.r, w.p, w.q, w.Minv, w.z, w.S, w.P);
               hipLaunchKernelGGL(k_cg_finish<CgFinBeta>, dim3(1), dim3(kFinBlock), 0, s, w.S, w.P, gu);
               hipLaunchKernelGGL(k_cg_pupdate, dim3(Blocks((n + 3) / 4)), dim3(256), 0, s, n, w.p, w.z, w.S);
            });
            if(t1 != 1 || t2 != 1) { (void)hipStreamSynchronize(s); (void)hipEventDestroy(ev); return -6; }
         }
         // read back the residual of the previous chunk while this one runs
         if(pending >= 0)
         {
            HIPCHK(hipEventSynchronize(ev));
            rr = w.host[1];
            if(rr <= goal) break;
         }
         HIPCHK(hipMemcpyAsync(&w.host[1], &w.S->rr, 8, hipMemcpyDeviceToHost, s));
         HIPCHK(hipEventRecord(ev, s));
         pending = it;
      }
      HIPCHK(hipStreamSynchronize(s));
      (void)hipEventDestroy(ev);
      return it;
   };

   int rc = 1;
   double rel = 0.;
   if(!refine)
   {
      int it = pcg((const float *)b->dev, tol, maxit);
      if(it < 0) { cleanup(); return it; }
      total = it;
      // FP64 residual of the FP32 solution
      hipLaunchKernelGGL(k_axpy64, dim3(nbn), dim3(256), 0, s, n, w.x64, w.x, 1.0);
   }
   else
   {
      for(int outer = 0; outer < 40 && total < maxit; outer++)
      {
         GML_B(k_resid64, dim3(gr), dim3(kCgBlock), 0, s, nrow, A->deg, A->off, A->col, (const float *)A->val,
               w.x64, (const float *)b->dev, w.r64, w.P);
         hipLaunchKernelGGL(k_cg_finish<CgFinResid>, dim3(1), dim3(kFinBlock), 0, s, w.S, w.P, gr);
         HIPCHK(hipMemcpyAsync(w.host, &w.S->rr0, 8, hipMemcpyDeviceToHost, s));
         HIPCHK(hipStreamSynchronize(s));
         double rn = sqrt(w.host[0]);
         rel = rn / bnorm;
         if(rel <= tol) break;
         // correction: A d = r / |r| in FP32
         hipLaunchKernelGGL(k_f64_to_f32, dim3(nbn), dim3(256), 0, s, n, w.r64, w.d32);
         hipLaunchKernelGGL(k_scale<float>, dim3(VecBlocks<float>(n)), dim3(256), 0, s, n, w.d32, (float)(1.0 / rn));
         int it = pcg(w.d32, 1e-4, maxit - total);
         if(it < 0) { cleanup(); return it; }
         total += it;
         hipLaunchKernelGGL(k_axpy64, dim3(nbn), dim3(256), 0, s, n, w.x64, w.x, rn);
      }
   }
   // the solution: FP64 as computed, or rounded to FP32
   if(xo->flt == GmlDbl)
      HIPCHK(hipMemcpyAsync(xo->dev, w.x64, n * 8, hipMemcpyDeviceToDevice, s));
   else
      hipLaunchKernelGGL(k_f64_to_f32, dim3(nbn), dim3(256), 0, s, n, w.x64, (float *)xo->dev);
   GML_B(k_resid64, dim3(gr), dim3(kCgBlock), 0, s, nrow, A->deg, A->off, A->col, (const float *)A->val,
         w.x64, (const float *)b->dev, w.r64, w.P);
   hipLaunchKernelGGL(k_cg_finish<CgFinResid>, dim3(1), dim3(kFinBlock), 0, s, w.S, w.P, gr);
   HIPCHK(hipMemcpyAsync(w.host, &w.S->rr0, 8, hipMemcpyDeviceToHost, s));
   if(hipStreamSynchronize(s) != hipSuccess) rc = -6;
   rel = sqrt(w.host[0]) / bnorm;
#undef GML_B
   *relres = rel;
   *iters = total;
   cleanup();
   return rc;
}

int GmlxDotVec(size_t GmlIdx, int VecA, int VecB, double *res)
{
   GML_CTX(g, GmlIdx);
   Vector *a = Vec(g, VecA), *b = Vec(g, VecB);
   if(!a || !b || a->flt != GmlFlt || b->flt != GmlFlt || a->NmbLin != b->NmbLin || a->b != b->b) return -1;
   if(g->host_only) return -6;
   CgScal *S = (CgScal *)DevAlloc(g, sizeof(CgScal));
   CgPart P = {(double *)DevAlloc(g, (size_t)2 * kCgGrid * 8), (unsigned int)kCgGrid};
   if(!S || !P.p) { DevFree(g, S); DevFree(g, P.p); return -5; }
   long long n = (long long)a->NmbLin * a->b;
   unsigned int gn = std::min<unsigned int>(kCgGrid, Blocks(n, kCgBlock));
   hipLaunchKernelGGL(k_dot, dim3(gn), dim3(kCgBlock), 0, g->stream, n, (const float *)a->dev, (const float *)b->dev, P);
   hipLaunchKernelGGL(k_cg_finish<CgFinDot>, dim3(1), dim3(kFinBlock), 0, g->stream, S, P, gn);
   bool ok = DevCopy(g, res, &S->pq, 8, 2);
   DevFree(g, S);
   DevFree(g, P.p);
   return ok ? 1 : -6;
}

} // extern "C"
