/*
 * gml_oracle.c -- CPU restatement of the GMlib kernel-engine semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke()
 * entry and the cpu_baseline leg of bench.py load it to compute expected
 * values.  The product (gmlib_amd/, libgml.so) never links, loads or calls it.
 *
 * Each function cites the reference lines whose behaviour it restates
 * (paths relative to the GMlib source tree).  Every function is pinned by
 * the CPU tests (tests/test_oracle_*.py, test_extract.py, test_bandwidth.py)
 * against closed forms, brute force on tiny meshes,
 * invariants and an FP64 restatement (there are no reference-held golden
 * vectors for this path: the reference ships no tests and its tet sample
 * meshes are missing).
 *
 * Built by oracle/Makefile into oracle/liboracle.so (gcc -O2 -fopenmp).
 * FP32 paths are compiled with -ffp-contract=off so that each operation is
 * rounded exactly as written.
 */

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { float x, y, z, w; } f4;
typedef struct { double x, y, z; } d3;

static f4 f4sub(f4 a, f4 b) { f4 r = {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; return r; }
static float f4dot(f4 a, f4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
static f4 f4cross(f4 a, f4 b)
{
   f4 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0.f};
   return r;
}

/* ------------------------------------------------------------------------ */
/* Topology: vertex -> element ball                                          */
/* ------------------------------------------------------------------------ */

/*
 * Ball of every vertex over an element table (nod_per_ele nodes per element,
 * row stride ele_stride ints).  Restates NewBallData for an uplink whose
 * source is the vertices (gmlib3.c:743-1168): elements are hashed in
 * ascending order with head insertion (AddHsh, gmlib3.c:1645-1674) and
 * fetched by walking the chain (GetHsh, gmlib3.c:1681-1721), so a vertex's
 * ball lists its elements in DESCENDING element index, each entry encoded as
 * (element << 4) | local_node_index (gmlib3.c:1706).
 *
 * deg[v] receives the degree; ball[v * width + k] the k-th entry for
 * k < min(deg, width), 0 elsewhere.  Returns the maximum degree.
 */
int orc_vertex_ball(int nver, int nele, int nod_per_ele, int ele_stride,
                    const int *ele, int width, int *deg, int *ball)
{
   int i, j, v, maxdeg = 0;
   int *fill;

   memset(deg, 0, sizeof(int) * (size_t)nver);
   if(ball)
      memset(ball, 0, sizeof(int) * (size_t)nver * (size_t)width);

   for(i = 0; i < nele; i++)
      for(j = 0; j < nod_per_ele; j++)
         deg[ ele[(size_t)i * ele_stride + j] ]++;

   for(v = 0; v < nver; v++)
      if(deg[v] > maxdeg)
         maxdeg = deg[v];

   if(!ball)
      return maxdeg;

   fill = calloc((size_t)nver, sizeof(int));

   /* walk elements from the highest index down: that is the chain order */
   for(i = nele - 1; i >= 0; i--)
      for(j = 0; j < nod_per_ele; j++)
      {
         v = ele[(size_t)i * ele_stride + j];
         if(fill[v] < width)
            ball[(size_t)v * width + fill[v]] = (i << 4) | j;
         fill[v]++;
      }

   free(fill);
   return maxdeg;
}

/*
 * Vector width rule of the generated code (GetCntVec, gmlib3.c:2945-2964):
 * p = ceil(log2 n) capped at 7; p <= 4 gives one vector of 2^p ints,
 * otherwise 2^(p-4) vectors of 16.
 */
void orc_get_cnt_vec(int n, int *cnt, int *vec)
{
   int p = (int)ceil(log2((double)n));
   if(p > 7) p = 7;
   if(p > 4) { *cnt = 1 << (p - 4); *vec = 16; }
   else      { *cnt = 1; *vec = 1 << p; }
}

/*
 * Numbering metric (GmlEvaluateNumbering, gmlib3.c:4409-4442): a 16-entry
 * direct-mapped tag table over 1024-index blocks, probed with the FIRST node
 * of every element (the reference loops j < NmbItm, which is 1 for element
 * tables).  Returns the hit percentage.
 */
float orc_evaluate_numbering(int nele, int ele_stride, const int *ele)
{
   int tag[64] = {0};
   long long hit = 0, mis = 0;
   int i, v, pos, adr;

   for(i = 0; i < nele; i++)
   {
      v = ele[(size_t)i * ele_stride];
      pos = (v >> 4) & 0xf;
      adr = v >> 10;
      if(tag[pos] == adr) hit++;
      else { mis++; tag[pos] = adr; }
   }
   return (float)(100 * (float)hit / (float)(hit + mis));
}

/* ------------------------------------------------------------------------ */
/* Geometry: quality formulas                                                */
/* ------------------------------------------------------------------------ */

/*
 * Mean-ratio tet quality, the per-tet kernel of the CPU-size quality
 * workload: q = 12 * cbrt(9 V^2) / sum(l_ij^2) * sign(V), V the signed
 * volume = ((b-a) x (c-a)).(d-a) / 6 (orientation as CalVol,
 * toolkit.cl:84-87).  1 for a regular tet, 0.7559526 for a Kuhn tet.
 * FP64 version: the accuracy reference.
 */
double orc_meanratio_f64(const double *a, const double *b, const double *c, const double *d)
{
   double ab[3], ac[3], ad[3], cr[3], v, s = 0.;
   const double *p[4] = {a, b, c, d};
   int i, j, k;

   for(k = 0; k < 3; k++) { ab[k] = b[k] - a[k]; ac[k] = c[k] - a[k]; ad[k] = d[k] - a[k]; }
   cr[0] = ab[1] * ac[2] - ab[2] * ac[1];
   cr[1] = ab[2] * ac[0] - ab[0] * ac[2];
   cr[2] = ab[0] * ac[1] - ab[1] * ac[0];
   v = (cr[0] * ad[0] + cr[1] * ad[1] + cr[2] * ad[2]) / 6.;

   for(i = 0; i < 4; i++)
      for(j = i + 1; j < 4; j++)
         for(k = 0; k < 3; k++)
            s += (p[j][k] - p[i][k]) * (p[j][k] - p[i][k]);

   return (v < 0 ? -1. : 1.) * 12. * cbrt(9. * v * v) / s;
}

/*
 * FP32 mean ratio in the exact operation order of the HIP kernels
 * (gmlib_amd/csrc/gml_mesh_kernels.hip, tet_meanratio): used for the
 * "same FP32 array" min checks and for FP32 round-off bounds.
 */
float orc_meanratio_f32(f4 a, f4 b, f4 c, f4 d)
{
   f4 e0 = f4sub(b, a), e1 = f4sub(c, a), e2 = f4sub(d, a);
   f4 e3 = f4sub(c, b), e4 = f4sub(d, b), e5 = f4sub(d, c);
   float s, v, q;
   e0.w = e1.w = e2.w = e3.w = e4.w = e5.w = 0.f;
   s = f4dot(e0, e0) + f4dot(e1, e1) + f4dot(e2, e2)
     + f4dot(e3, e3) + f4dot(e4, e4) + f4dot(e5, e5);
   v = f4dot(f4cross(e0, e1), e2) * (1.f / 6.f);
   q = 12.f * cbrtf(9.f * v * v) / s;
   return v < 0.f ? -q : q;
}

/*
 * Reference "compute_quality" body semantics (examples/MeshQuality/
 * compute_quality.cl:1-17): len = max squared edge length; srf = area of
 * faces (0,1,2) and (0,1,3) only -- the stray ';' on line 11 turns lines
 * 12-13 into no-ops; vol = .166666 * CalVol; qal = 7.348469 * vol /
 * (srf * sqrt(len)).  CalTriSrf = .5 * |(c-a) x (b-a)| (toolkit.cl:79-82,
 * 120-123).  FP64 evaluation; a regular tet gives 0.999996.
 */
double orc_ref_compute_quality_f64(const double *a, const double *b, const double *c, const double *d)
{
   const double *p[4] = {a, b, c, d};
   double len = 0., srf, vol, t, u[3], w[3], cr[3];
   int i, j, k;
   int fac[2][3] = {{0, 1, 2}, {0, 1, 3}};

   for(i = 0; i < 4; i++)
      for(j = i + 1; j < 4; j++)
      {
         t = 0.;
         for(k = 0; k < 3; k++) t += (p[i][k] - p[j][k]) * (p[i][k] - p[j][k]);
         if(t > len) len = t;
      }

   srf = 0.;
   for(i = 0; i < 2; i++)
   {
      const double *fa = p[fac[i][0]], *fb = p[fac[i][1]], *fc = p[fac[i][2]];
      for(k = 0; k < 3; k++) { u[k] = fc[k] - fa[k]; w[k] = fb[k] - fa[k]; }
      cr[0] = u[1] * w[2] - u[2] * w[1];
      cr[1] = u[2] * w[0] - u[0] * w[2];
      cr[2] = u[0] * w[1] - u[1] * w[0];
      srf += .5 * sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
   }

   for(k = 0; k < 3; k++) { u[k] = b[k] - a[k]; w[k] = c[k] - a[k]; }
   cr[0] = u[1] * w[2] - u[2] * w[1];
   cr[1] = u[2] * w[0] - u[0] * w[2];
   cr[2] = u[0] * w[1] - u[1] * w[0];
   vol = .166666 * (cr[0] * (d[0] - a[0]) + cr[1] * (d[1] - a[1]) + cr[2] * (d[2] - a[2]));

   return 7.348469 * vol / (srf * sqrt(len));
}

/*
 * Toolkit CalTetQal (toolkit.cl:209-228), FP64: h sums CalLen(b,c) twice and
 * omits (b,d) (kept as written); s = sum of the four |cross| face terms
 * (CalSrf = 2 * area); v = CalVol = 6 * volume; q = 176.363 * v / (h * s).
 */
double orc_caltetqal_f64(const double *a, const double *b, const double *c, const double *d)
{
   const double *P[4] = {a, b, c, d};
   double h = 0., s = 0., v, u[3], w[3], cr[3];
   int e[6][2] = {{0,1},{0,2},{0,3},{1,2},{1,2},{2,3}};
   int f[4][3] = {{0,1,2},{2,0,3},{1,2,3},{3,0,1}};
   int i, k;

   for(i = 0; i < 6; i++)
   {
      double t = 0.;
      for(k = 0; k < 3; k++) t += (P[e[i][1]][k] - P[e[i][0]][k]) * (P[e[i][1]][k] - P[e[i][0]][k]);
      h += sqrt(t);
   }
   for(i = 0; i < 4; i++)
   {
      const double *fa = P[f[i][0]], *fb = P[f[i][1]], *fc = P[f[i][2]];
      for(k = 0; k < 3; k++) { u[k] = fc[k] - fa[k]; w[k] = fb[k] - fa[k]; }
      cr[0] = u[1] * w[2] - u[2] * w[1];
      cr[1] = u[2] * w[0] - u[0] * w[2];
      cr[2] = u[0] * w[1] - u[1] * w[0];
      s += sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
   }
   for(k = 0; k < 3; k++) { u[k] = b[k] - a[k]; w[k] = c[k] - a[k]; }
   cr[0] = u[1] * w[2] - u[2] * w[1];
   cr[1] = u[2] * w[0] - u[0] * w[2];
   cr[2] = u[0] * w[1] - u[1] * w[0];
   v = cr[0] * (d[0] - a[0]) + cr[1] * (d[1] - a[1]) + cr[2] * (d[2] - a[2]);
   return 176.363 * v / (h * s);
}

/* Triangle quality CalTriQal (toolkit.cl:166-178), FP64. */
double orc_caltriqal_f64(const double *a, const double *b, const double *c)
{
   double ha = 0., hb = 0., hc = 0., hmax, u[3], w[3], cr[3];
   int k;
   for(k = 0; k < 3; k++)
   {
      ha += (c[k] - b[k]) * (c[k] - b[k]);
      hb += (a[k] - c[k]) * (a[k] - c[k]);
      hc += (b[k] - a[k]) * (b[k] - a[k]);
      u[k] = c[k] - a[k]; w[k] = b[k] - a[k];
   }
   ha = sqrt(ha); hb = sqrt(hb); hc = sqrt(hc);
   hmax = ha > hb ? ha : hb;
   hmax = hmax > hc ? hmax : hc;
   cr[0] = u[1] * w[2] - u[2] * w[1];
   cr[1] = u[2] * w[0] - u[0] * w[2];
   cr[2] = u[0] * w[1] - u[1] * w[0];
   return 3.46410 * sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]) / (hmax * (ha + hb + hc));
}

/* Mean-ratio quality of every tet, FP64 from FP32 coordinates (crd: float4 rows). */
void orc_tet_meanratio_f64_all(int ntet, const int *tet, const float *crd, double *q)
{
   int i;
#pragma omp parallel for schedule(static)
   for(i = 0; i < ntet; i++)
   {
      double p[4][3];
      int j, k;
      for(j = 0; j < 4; j++)
         for(k = 0; k < 3; k++)
            p[j][k] = crd[(size_t)tet[(size_t)i * 4 + j] * 4 + k];
      q[i] = orc_meanratio_f64(p[0], p[1], p[2], p[3]);
   }
}

/* Same in the FP32 kernel order. */
void orc_tet_meanratio_f32_all(int ntet, const int *tet, const float *crd, float *q)
{
   int i;
#pragma omp parallel for schedule(static)
   for(i = 0; i < ntet; i++)
   {
      const f4 *c4 = (const f4 *)crd;
      const int *t = &tet[(size_t)i * 4];
      q[i] = orc_meanratio_f32(c4[t[0]], c4[t[1]], c4[t[2]], c4[t[3]]);
   }
}

/* ------------------------------------------------------------------------ */
/* Ball gathers                                                              */
/* ------------------------------------------------------------------------ */

/*
 * Per-vertex minimum of an element field over the vertex ball (the uplink
 * gather of examples/BallQuality/verqal.cl:1-6 with a min body): out[v] =
 * min over the deg[v] ball entries of val[entry >> 4]; a vertex with an
 * empty ball gets 0 (the generated code reads the null value for slot 0).
 * Exact (min is exact in any order).
 */
void orc_ball_min(int nver, const int *deg, const int *ball, int width,
                  const float *val, float *out)
{
   int v;
#pragma omp parallel for schedule(static)
   for(v = 0; v < nver; v++)
   {
      float m = deg[v] > 0 ? FLT_MAX : 0.f;
      int k;
      for(k = 0; k < deg[v] && k < width; k++)
      {
         float x = val[ ball[(size_t)v * width + k] >> 4 ];
         m = x < m ? x : m;
      }
      out[v] = m;
   }
}

/* ------------------------------------------------------------------------ */
/* Reductions                                                                */
/* ------------------------------------------------------------------------ */

/*
 * GmlReduceVector operation meanings (gmlib3.c:3229-3332, reduce.cl:6-137):
 * 0 min, 1 max, 2 sum, 3 L0 (count of nonzeros), 4 L1 (sum |x|), 5 L2
 * (sum x^2, no square root), 6 Linf (max |x|).  Evaluated exactly over the
 * whole vector in FP64 (the reference finishes in float and drops the last
 * partial group; the drop-in replacement reduces the full vector).
 * min/max/Linf are exact; the sums use FP64 with Kahan compensation.
 */
double orc_reduce(int op, long long n, const float *x)
{
   long long i;
   double r, c = 0., y, t;

   switch(op)
   {
      case 0: { float m = FLT_MAX;  for(i = 0; i < n; i++) m = fminf(m, x[i]); return m; }
      case 1: { float m = -FLT_MAX; for(i = 0; i < n; i++) m = fmaxf(m, x[i]); return m; }
      case 6: { float m = 0.f;      for(i = 0; i < n; i++) m = fmaxf(m, fabsf(x[i])); return m; }
      case 3: { long long k = 0;    for(i = 0; i < n; i++) k += (x[i] != 0.f); return (double)k; }
      default: break;
   }

   r = 0.;
   for(i = 0; i < n; i++)
   {
      double v = x[i];
      if(op == 4) v = fabs(v);
      else if(op == 5) v = v * v;
      y = v - c;
      t = r + y;
      c = (t - r) - y;
      r = t;
   }
   return r;
}

/* ------------------------------------------------------------------------ */
/* Node smoothing                                                            */
/* ------------------------------------------------------------------------ */

/*
 * One Jacobi sweep of barycentre Laplacian smoothing, the scale workload of
 * this build, written as the two GMlib kernels of the NodeSmoothing pattern
 * (examples/NodeSmoothing/NodeSmoothing.c:137-161):
 *   scatter (per tet):   bary[t] = (((p0 + p1) + p2) + p3) * 0.25
 *   gather  (per vertex, ball in reference order, GmlRefFlag on VerCrd):
 *       s = sum over k < deg of bary[ball[k] >> 4]   (in ball order)
 *       avg = s * (1 / deg)
 *       ref != 0 (boundary): unchanged; else p = p + (avg - p) * omega
 * FP32, operations rounded as written (no contraction).  crd is float4
 * rows, updated in place; bary is caller scratch of ntet float4.
 */
void orc_smooth_sweep_f32(int nver, int ntet, const int *tet, const int *ref,
                          const int *deg, const int *ball, int width,
                          float *crd, float *bary, float omega)
{
   int i;
   f4 *c4 = (f4 *)crd, *b4 = (f4 *)bary;

#pragma omp parallel for schedule(static)
   for(i = 0; i < ntet; i++)
   {
      const int *t = &tet[(size_t)i * 4];
      f4 p0 = c4[t[0]], p1 = c4[t[1]], p2 = c4[t[2]], p3 = c4[t[3]], r;
      r.x = (((p0.x + p1.x) + p2.x) + p3.x) * 0.25f;
      r.y = (((p0.y + p1.y) + p2.y) + p3.y) * 0.25f;
      r.z = (((p0.z + p1.z) + p2.z) + p3.z) * 0.25f;
      r.w = (((p0.w + p1.w) + p2.w) + p3.w) * 0.25f;
      b4[i] = r;
   }

#pragma omp parallel for schedule(static)
   for(i = 0; i < nver; i++)
   {
      f4 s = {0.f, 0.f, 0.f, 0.f}, p = c4[i];
      float inv;
      int k, d = deg[i];

      if(ref[i] || d == 0)
         continue;

      for(k = 0; k < d; k++)
      {
         f4 b = b4[ ball[(size_t)i * width + k] >> 4 ];
         s.x += b.x; s.y += b.y; s.z += b.z; s.w += b.w;
      }
      inv = 1.f / (float)d;
      s.x *= inv; s.y *= inv; s.z *= inv; s.w *= inv;
      p.x = p.x + (s.x - p.x) * omega;
      p.y = p.y + (s.y - p.y) * omega;
      p.z = p.z + (s.z - p.z) * omega;
      c4[i] = p;
   }
}

/*
 * FP64 restatement of the same sweep (the accuracy oracle): coordinates held
 * as double triples (crd: nver*3), barycentres recomputed in FP64.  The ball
 * order does not matter beyond round-off here.
 */
void orc_smooth_sweep_f64(int nver, int ntet, const int *tet, const int *ref,
                          const int *deg, const int *ball, int width,
                          double *crd, double *bary, double omega)
{
   int i;

#pragma omp parallel for schedule(static)
   for(i = 0; i < ntet; i++)
   {
      const int *t = &tet[(size_t)i * 4];
      int k;
      for(k = 0; k < 3; k++)
         bary[(size_t)i * 3 + k] = (crd[(size_t)t[0] * 3 + k] + crd[(size_t)t[1] * 3 + k]
                                  + crd[(size_t)t[2] * 3 + k] + crd[(size_t)t[3] * 3 + k]) * 0.25;
   }

#pragma omp parallel for schedule(static)
   for(i = 0; i < nver; i++)
   {
      double s[3] = {0., 0., 0.};
      int k, j, d = deg[i];
      if(ref[i] || d == 0)
         continue;
      for(k = 0; k < d; k++)
      {
         int e = ball[(size_t)i * width + k] >> 4;
         for(j = 0; j < 3; j++) s[j] += bary[(size_t)e * 3 + j];
      }
      for(j = 0; j < 3; j++)
         crd[(size_t)i * 3 + j] += (s[j] / d - crd[(size_t)i * 3 + j]) * omega;
   }
}

/*
 * Reference NodeSmoothing sweep (examples/NodeSmoothing/scatter.cl:1-46 and
 * gather.cl:1-17), FP64: per tet and per face, the ideal apex = face
 * centroid + 0.816496 * sqrt(max squared edge of the face) * unit normal
 * (normal = normalize(cross) with the edge pairs of scatter.cl); per vertex,
 * the mean of the ideal positions selected by the voyeur (local index) over
 * the ball, relaxed with mix(p, mean, 0.2); res[v] = |p - new|^2.
 * Used to pin the generic-codegen path that runs those bodies.
 */
void orc_ref_nodesmooth_sweep_f64(int nver, int ntet, const int *tet,
                                  const int *deg, const int *ball, int width,
                                  double *crd, double *opt, double *res)
{
   int i;
   /* edges e0..e5 = p0-p1, p0-p2, p0-p3, p1-p2, p1-p3, p2-p3 */
   static const int eidx[6][2] = {{0,1},{0,2},{0,3},{1,2},{1,3},{2,3}};
   /* per opposite vertex f: the three edges whose max defines the height,
      the triangle vertices of the centroid and the cross-product pair */
   static const int emax[4][3] = {{3,4,5},{1,2,5},{0,2,4},{0,1,3}};
   static const int ctr[4][3]  = {{1,2,3},{0,2,3},{1,0,3},{1,2,0}};
   static const int crs[4][2]  = {{4,5},{5,2},{2,4},{3,1}};

#pragma omp parallel for schedule(static)
   for(i = 0; i < ntet; i++)
   {
      double P[4][3], E[6][3], L[6];
      int j, k, f;
      for(j = 0; j < 4; j++)
         for(k = 0; k < 3; k++)
            P[j][k] = crd[(size_t)tet[(size_t)i * 4 + j] * 3 + k];
      for(j = 0; j < 6; j++)
      {
         L[j] = 0.;
         for(k = 0; k < 3; k++)
         {
            E[j][k] = P[eidx[j][0]][k] - P[eidx[j][1]][k];
            L[j] += E[j][k] * E[j][k];
         }
      }
      for(f = 0; f < 4; f++)
      {
         double m = L[emax[f][0]], c[3], n[3], nl;
         const double *a = E[crs[f][0]], *b = E[crs[f][1]];
         if(L[emax[f][1]] > m) m = L[emax[f][1]];
         if(L[emax[f][2]] > m) m = L[emax[f][2]];
         m = 0.816496 * sqrt(m);
         n[0] = a[1] * b[2] - a[2] * b[1];
         n[1] = a[2] * b[0] - a[0] * b[2];
         n[2] = a[0] * b[1] - a[1] * b[0];
         nl = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
         for(k = 0; k < 3; k++)
         {
            c[k] = (P[ctr[f][0]][k] + P[ctr[f][1]][k] + P[ctr[f][2]][k]) * (1. / 3.);
            opt[((size_t)i * 4 + f) * 3 + k] = c[k] + n[k] / nl * m;
         }
      }
   }

#pragma omp parallel for schedule(static)
   for(i = 0; i < nver; i++)
   {
      double s[3] = {0., 0., 0.}, nw[3], r = 0.;
      int k, j, d = deg[i];
      for(k = 0; k < d; k++)
      {
         int e = ball[(size_t)i * width + k];
         for(j = 0; j < 3; j++) s[j] += opt[((size_t)(e >> 4) * 4 + (e & 7)) * 3 + j];
      }
      for(j = 0; j < 3; j++)
      {
         double a = d ? s[j] / d : 0.;
         nw[j] = crd[(size_t)i * 3 + j] + (a - crd[(size_t)i * 3 + j]) * 0.2;
         r += (crd[(size_t)i * 3 + j] - nw[j]) * (crd[(size_t)i * 3 + j] - nw[j]);
         crd[(size_t)i * 3 + j] = nw[j];
      }
      res[i] = r;
   }
}

/*
 * Volume-weighted opposite-face pull, FP64: the two-body uplink pattern the
 * codegen tests run (a tet body writing 4 float4 items per tet, a vertex
 * body reading them through the vertex -> tet ball with voyeurs, the way
 * examples/NodeSmoothing/NodeSmoothing.c:137-141 wires its bodies).  Per tet
 * i and local vertex j: item (i, j) = (centroid of the face opposite j,
 * w = |volume|).  Per vertex with a non-empty ball: target = sum over its
 * ball entries (i, j) of w_i * centroid_ij / sum of w_i; new = p + 0.25 *
 * (target - p); res = |new - p|^2.  An empty ball leaves p and res = 0.
 * crd: nver * 3 doubles (updated), item: ntet * 4 * 4 doubles.
 */
void orc_face_pull_sweep_f64(int nver, int ntet, const int *tet, const int *deg, const int *ball, int width,
                             double *crd, double *item, double *res)
{
   int i;
#pragma omp parallel for schedule(static)
   for(i = 0; i < ntet; i++)
   {
      double P[4][3], s[3] = {0., 0., 0.}, u[3], v[3], w[3], vol;
      int j, k;
      for(j = 0; j < 4; j++)
         for(k = 0; k < 3; k++)
         {
            P[j][k] = crd[(size_t)tet[(size_t)i * 4 + j] * 3 + k];
            s[k] += P[j][k];
         }
      for(k = 0; k < 3; k++) { u[k] = P[1][k] - P[0][k]; v[k] = P[2][k] - P[0][k]; w[k] = P[3][k] - P[0][k]; }
      vol = fabs((u[1] * v[2] - u[2] * v[1]) * w[0] + (u[2] * v[0] - u[0] * v[2]) * w[1]
               + (u[0] * v[1] - u[1] * v[0]) * w[2]) / 6.;
      for(j = 0; j < 4; j++)
      {
         for(k = 0; k < 3; k++) item[((size_t)i * 4 + j) * 4 + k] = (s[k] - P[j][k]) / 3.;
         item[((size_t)i * 4 + j) * 4 + 3] = vol;
      }
   }
#pragma omp parallel for schedule(static)
   for(i = 0; i < nver; i++)
   {
      double a[3] = {0., 0., 0.}, ws = 0., r = 0.;
      int k, j, d = deg[i];
      for(k = 0; k < d; k++)
      {
         const int e = ball[(size_t)i * width + k];
         const double *it = &item[((size_t)(e >> 4) * 4 + (e & 7)) * 4];
         for(j = 0; j < 3; j++) a[j] += it[3] * it[j];
         ws += it[3];
      }
      if(d > 0)
         for(j = 0; j < 3; j++)
         {
            const double p = crd[(size_t)i * 3 + j], n = p + (a[j] / ws - p) * 0.25;
            r += (n - p) * (n - p);
            crd[(size_t)i * 3 + j] = n;
         }
      res[i] = r;
   }
}

/* ------------------------------------------------------------------------ */
/* Block sparse algebra                                                      */
/* ------------------------------------------------------------------------ */

/*
 * y = A x for a block-CSR matrix (b x b row-major blocks, FP32 values as the
 * GmlNewMatrix input, gmlib3.c:1175-1427 and multmatvec.cl:30-67): the
 * products accumulated in FP64.  x, y: nrow * b doubles.
 */
void orc_bsr_spmv_f64(int nrow, int b, const int *rowptr, const int *col,
                      const float *val, const double *x, double *y)
{
   int r;
#pragma omp parallel for schedule(static)
   for(r = 0; r < nrow; r++)
   {
      double acc[16] = {0};
      int k, i, j;
      for(k = rowptr[r]; k < rowptr[r + 1]; k++)
      {
         const float *blk = &val[(size_t)k * b * b];
         const double *xc = &x[(size_t)col[k] * b];
         for(i = 0; i < b; i++)
            for(j = 0; j < b; j++)
               acc[i] += (double)blk[i * b + j] * xc[j];
      }
      for(i = 0; i < b; i++)
         y[(size_t)r * b + i] = acc[i];
   }
}

/* Same product with FP32 values, vectors and accumulation (row order as written). */
void orc_bsr_spmv_f32(int nrow, int b, const int *rowptr, const int *col,
                      const float *val, const float *x, float *y)
{
   int r;
#pragma omp parallel for schedule(static)
   for(r = 0; r < nrow; r++)
   {
      float acc[16] = {0};
      int k, i, j;
      for(k = rowptr[r]; k < rowptr[r + 1]; k++)
      {
         const float *blk = &val[(size_t)k * b * b];
         const float *xc = &x[(size_t)col[k] * b];
         for(i = 0; i < b; i++)
         {
            float t = 0.f;
            for(j = 0; j < b; j++)
               t += blk[i * b + j] * xc[j];
            acc[i] += t;
         }
      }
      for(i = 0; i < b; i++)
         y[(size_t)r * b + i] = acc[i];
   }
}

/* ------------------------------------------------------------------------ */
/* Topology: edge and face extraction                                        */
/* ------------------------------------------------------------------------ */

/*
 * Distinct items (edges: deg 2, triangles: 3, quads: 4) of a list of
 * element sources, restating the hash of GmlExtractFaces
 * (gmlib3.c:3880-4171) with its helpers: each item's nodes sorted ascending
 * (GetItmNod, gmlib3.c:1567-1620), hash key sum(w_i * node_i) mod TabSiz
 * with w = {3, 5, 7, 11} and TabSiz = the number of items hashed
 * (CalHshKey, gmlib3.c:1627-1638), buckets chained by head insertion
 * (AddHsh, gmlib3.c:1645-1674) and searched by node comparison along the
 * chain (GetHsh, gmlib3.c:1681-1721), items visited source by source,
 * element by element, local item by local item.
 *
 * The reference writes the output in a second pass: an existing item i is
 * written when its first hashed appearance is i itself (gmlib3.c:4020-4035),
 * an element item when its first appearance is (this kind, this element)
 * (gmlib3.c:4089-4100).  With one appearance of an item per element, that is
 * the list of first appearances in visiting order, which is what this
 * function emits directly when an item is first added.  Source 0 may be the
 * existing table (has_old): its items keep the ref of their first
 * occurrence (oldref), all other items get ref 0.
 *
 * Sources: ne[s] elements of stride[s] ints at ele[s]; nloc[s] local items,
 * local node indices loc[(s * 12 + j) * 4 + k].  tabsiz: the number of hash
 * heads, 0 for the number of items hashed (the face tables' size); the edge
 * table of the reference has one head per element (gmlib3.c:3730-3736).
 * out_nodes (deg ints per item) and out_ref receive at most cap items;
 * returns the item count.
 *
 * Edges: the reference's GmlExtractEdges (gmlib3.c:3719-3873) does not hash
 * the existing edge table, which makes its handling of existing edges write
 * past the new table; the face rule above (existing items hashed first) is
 * applied to edges too.
 */
long long orc_extract_items(int deg, int nsrc, const long long *ne, const int *stride, const int *nloc,
                            const int *loc, const int *const *ele, int has_old, const int *oldref,
                            long long tabsiz, int *out_nodes, int *out_ref, long long cap)
{
   static const int wei[4] = {3, 5, 7, 11};
   typedef struct { int nod[4], nxt, ref; } buc_t;
   long long tab = 0, nbuc = 1, nout = 0, s, e;
   int *head;
   buc_t *buc;

   for(s = 0; s < nsrc; s++)
      tab += ne[s] * nloc[s];
   if(!tab)
      return 0;
   if(tabsiz > 0)
      tab = tabsiz;
   head = calloc((size_t)tab, sizeof(int));
   for(s = 0, nbuc = 0; s < nsrc; s++)
      nbuc += ne[s] * nloc[s];
   buc = malloc((size_t)(nbuc + 1) * sizeof(buc_t));
   nbuc = 1;
   if(!head || !buc)
   {
      free(head);
      free(buc);
      return -1;
   }

   for(s = 0; s < nsrc; s++)
      for(e = 0; e < ne[s]; e++)
      {
         const int *en = ele[s] + e * stride[s];
         int j;
         for(j = 0; j < nloc[s]; j++)
         {
            int n[4] = {0, 0, 0, 0}, i, k, key = 0, b, hit = 0;
            for(i = 0; i < deg; i++)
               n[i] = en[ loc[(s * 12 + j) * 4 + i] ];
            for(i = 0; i < deg; i++)
               for(k = i + 1; k < deg; k++)
                  if(n[i] > n[k]) { int t = n[i]; n[i] = n[k]; n[k] = t; }
            for(i = 0; i < deg; i++)
               key += wei[i] * n[i];
            key = (int)(key % tab);
            for(b = head[key]; b && !hit; b = buc[b].nxt)
               hit = !memcmp(buc[b].nod, n, sizeof(n));
            if(hit)
               continue;
            memcpy(buc[nbuc].nod, n, sizeof(n));
            buc[nbuc].nxt = head[key];
            buc[nbuc].ref = (has_old && s == 0 && oldref) ? oldref[e] : 0;
            head[key] = (int)nbuc;
            if(nout < cap)
            {
               memcpy(&out_nodes[nout * deg], n, (size_t)deg * sizeof(int));
               out_ref[nout] = buc[nbuc].ref;
            }
            nbuc++;
            nout++;
         }
      }
   free(head);
   free(buc);
   return nout;
}

/* ------------------------------------------------------------------------ */
/* GpuBandwidth access patterns                                              */
/* ------------------------------------------------------------------------ */

/*
 * The bodies of examples/GpuBandwidth (driven by gpu_bandwidth.c:108-176),
 * with the generated prologue's reads: direct_tets.cl converts each node
 * index of a tet to float (TetDat = (float4){TetVer.s0..s3}); indirect_tets.cl
 * sums the tet's four vertex coordinates left to right through the element
 * table (VerCrd[k] = VerCrdTab[TetVer.sk], gmlib3.c:2777-2880); vertices_ball.cl
 * sums TetDat over the vertex ball in ball order, slots past the degree
 * reading the null value 0 (adding +0 leaves a sum that started at +0
 * unchanged, so they are skipped here).  FP32, each addition rounded as
 * written (-ffp-contract=off).
 */
void orc_bw_direct_tets(int ntet, const int *tet, float *out)
{
   long long i;
#pragma omp parallel for schedule(static)
   for(i = 0; i < 4LL * ntet; i++)
      out[i] = (float)tet[i];
}

void orc_bw_indirect_tets(int ntet, const int *tet, const float *crd, float *out)
{
   int i;
#pragma omp parallel for schedule(static)
   for(i = 0; i < ntet; i++)
   {
      const int *t = &tet[(size_t)i * 4];
      int c;
      for(c = 0; c < 4; c++)
         out[(size_t)i * 4 + c] = ((crd[(size_t)t[0] * 4 + c] + crd[(size_t)t[1] * 4 + c]) + crd[(size_t)t[2] * 4 + c])
                                  + crd[(size_t)t[3] * 4 + c];
   }
}

void orc_bw_vertices_ball(int nver, const int *deg, const int *ball, int width, const float *tdat, float *out)
{
   int v;
#pragma omp parallel for schedule(static)
   for(v = 0; v < nver; v++)
   {
      float a[4] = {0.f, 0.f, 0.f, 0.f};
      int k, c;
      for(k = 0; k < deg[v] && k < width; k++)
      {
         const float *x = &tdat[(size_t)(ball[(size_t)v * width + k] >> 4) * 4];
         for(c = 0; c < 4; c++) a[c] += x[c];
      }
      for(c = 0; c < 4; c++) out[(size_t)v * 4 + c] = a[c];
   }
}

/* Thread count the OpenMP regions above use (for the CPU baseline record). */
int orc_num_threads(void)
{
#ifdef _OPENMP
   return omp_get_max_threads();
#else
   return 1;
#endif
}
