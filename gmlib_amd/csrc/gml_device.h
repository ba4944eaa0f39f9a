// Device-side helpers shared by the hand-written mesh kernels
// (gml_mesh_kernels.hip, gml_batch.hip).

#pragma once

#include <hip/hip_runtime.h>

namespace gml {

// clang ext-vector types: arrays of these stay in registers (HIP's float4 is
// a struct wrapper whose arrays the compiler may place in scratch)
typedef float v4f __attribute__((ext_vector_type(4)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef unsigned short v4u16 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4i ld_nt(const v4i *p) { return __builtin_nontemporal_load(p); }

__device__ __forceinline__ float dot4(v4f a, v4f b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }

// mean ratio: the expression sequence of tet_meanratio.cl
__device__ __forceinline__ float meanratio(v4f p0, v4f p1, v4f p2, v4f p3)
{
   v4f e0 = p1 - p0, e1 = p2 - p0, e2 = p3 - p0, e3 = p2 - p1, e4 = p3 - p1, e5 = p3 - p2;
   float s = dot4(e0, e0) + dot4(e1, e1) + dot4(e2, e2) + dot4(e3, e3) + dot4(e4, e4) + dot4(e5, e5);
   v4f c = {e0.y * e1.z - e0.z * e1.y, e0.z * e1.x - e0.x * e1.z, e0.x * e1.y - e0.y * e1.x, 0.f};
   float v = dot4(c, e2) * (1.f / 6.f);
   float q = 12.f * cbrtf(9.f * v * v) / s;
   if(v < 0.f) q = -q;
   return q;
}

__device__ __forceinline__ float dot3(v4f a, v4f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// The same mean ratio for coordinates whose fourth component is 0: every
// term dropped from meanratio() is then an exact 0 added last (fma(0, 0, a)
// = a), so the result is bit-identical and costs fewer instructions.
__device__ __forceinline__ float meanratio3(v4f p0, v4f p1, v4f p2, v4f p3)
{
   v4f e0 = p1 - p0, e1 = p2 - p0, e2 = p3 - p0, e3 = p2 - p1, e4 = p3 - p1, e5 = p3 - p2;
   float s = dot3(e0, e0) + dot3(e1, e1) + dot3(e2, e2) + dot3(e3, e3) + dot3(e4, e4) + dot3(e5, e5);
   v4f c = {e0.y * e1.z - e0.z * e1.y, e0.z * e1.x - e0.x * e1.z, e0.x * e1.y - e0.y * e1.x, 0.f};
   float v = dot3(c, e2) * (1.f / 6.f);
   float q = 12.f * cbrtf(9.f * v * v) / s;
   if(v < 0.f) q = -q;
   return q;
}

} // namespace gml
