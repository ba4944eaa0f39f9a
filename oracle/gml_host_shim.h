// Host (x86-64) stand-ins for the HIP qualifiers and builtins that the
// engine's generated kernels use, so that a generated kernel's text --
// the OpenCL-C prelude, the GMlib prologue/epilogue the engine writes and
// the project's OpenCL-C body -- compiles unchanged with clang for the CPU.
//
// Measurement and test infrastructure only (the CPU baseline): nothing in
// the product includes this file.  A work-item is one call of the kernel
// function with blockIdx / threadIdx set for it; host_kernels.py writes the
// loop that makes those calls over the host cores.  Only the one-row launch
// form without LDS staging is supported: it has no work-group barriers, so
// work-items can run one after another.
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline __attribute__((always_inline))
#define __launch_bounds__(...)
#define __shared__ __shared__is_not_supported_by_the_host_shim

struct gml_host_dim { unsigned int x, y, z; };
// the work-item being run by this host thread, as a HIP kernel sees it
static thread_local gml_host_dim blockIdx, threadIdx, blockDim, gridDim;

// The engine marks streamed rows nontemporal for the GPU's caches; on x86 a
// nontemporal store evicts the line the work-item has just read, which makes
// the gather several times slower, and the reference's OpenCL kernels carry
// no such hint, so the host build uses ordinary loads and stores.
#define __builtin_nontemporal_load(p) (*(p))
#define __builtin_nontemporal_store(v, p) (*(p) = (v))

static inline void __syncthreads() {}
static inline void __threadfence_block() {}
static inline float rsqrtf(float x) { return 1.f / sqrtf(x); }
static inline double rsqrt(double x) { return 1. / sqrt(x); }
static inline float __int_as_float(int x) { float f; memcpy(&f, &x, 4); return f; }
static inline int __float_as_int(float f) { int x; memcpy(&x, &f, 4); return x; }

// HIP's atomics on host memory (relaxed, as the device's)
template <class T> static inline T atomicAdd(T *p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }
template <class T> static inline T atomicSub(T *p, T v) { return __atomic_fetch_sub(p, v, __ATOMIC_RELAXED); }
template <class T> static inline T atomicAnd(T *p, T v) { return __atomic_fetch_and(p, v, __ATOMIC_RELAXED); }
template <class T> static inline T atomicOr(T *p, T v) { return __atomic_fetch_or(p, v, __ATOMIC_RELAXED); }
template <class T> static inline T atomicXor(T *p, T v) { return __atomic_fetch_xor(p, v, __ATOMIC_RELAXED); }
template <class T> static inline T atomicExch(T *p, T v) { return __atomic_exchange_n(p, v, __ATOMIC_RELAXED); }
static inline float atomicExch(float *p, float v)
{
   int o = __atomic_exchange_n((int *)p, __float_as_int(v), __ATOMIC_RELAXED);
   return __int_as_float(o);
}
template <class T> static inline T atomicCAS(T *p, T c, T v)
{
   __atomic_compare_exchange_n(p, &c, v, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED);
   return c;
}
template <class T> static inline T atomicMin(T *p, T v)
{
   T o = __atomic_load_n(p, __ATOMIC_RELAXED);
   while(v < o && !__atomic_compare_exchange_n(p, &o, v, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {}
   return o;
}
template <class T> static inline T atomicMax(T *p, T v)
{
   T o = __atomic_load_n(p, __ATOMIC_RELAXED);
   while(v > o && !__atomic_compare_exchange_n(p, &o, v, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {}
   return o;
}

// HIP's device overloads of min / max for scalars
static inline int min(int a, int b) { return a < b ? a : b; }
static inline int max(int a, int b) { return a > b ? a : b; }
static inline unsigned int min(unsigned int a, unsigned int b) { return a < b ? a : b; }
static inline unsigned int max(unsigned int a, unsigned int b) { return a > b ? a : b; }
static inline long min(long a, long b) { return a < b ? a : b; }
static inline long max(long a, long b) { return a > b ? a : b; }
static inline float min(float a, float b) { return fminf(a, b); }
static inline float max(float a, float b) { return fmaxf(a, b); }
static inline double min(double a, double b) { return fmin(a, b); }
static inline double max(double a, double b) { return fmax(a, b); }
