// 3D Hilbert index of integer coordinates (J. Skilling, "Programming the
// Hilbert curve", AIP Conf. Proc. 707, 381 (2004): axes to transposed
// index, then the bits interleaved x, y, z from the most significant).
// Host and device.

#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace gml {

__host__ __device__ inline uint64_t HilbertKey(unsigned x, unsigned y, unsigned z, int bits)
{
   unsigned X[3] = {x, y, z};
   unsigned M = 1u << (bits - 1);
   for(unsigned Q = M; Q > 1; Q >>= 1)
   {
      unsigned P = Q - 1;
      for(int i = 0; i < 3; i++)
      {
         if(X[i] & Q) X[0] ^= P;
         else { unsigned t = (X[0] ^ X[i]) & P; X[0] ^= t; X[i] ^= t; }
      }
   }
   for(int i = 1; i < 3; i++) X[i] ^= X[i - 1];
   unsigned t = 0;
   for(unsigned Q = M; Q > 1; Q >>= 1)
      if(X[2] & Q) t ^= Q - 1;
   for(int i = 0; i < 3; i++) X[i] ^= t;
   uint64_t h = 0;
   for(int b = bits - 1; b >= 0; b--)
      for(int i = 0; i < 3; i++)
         h = (h << 1) | ((X[i] >> b) & 1u);
   return h;
}

} // namespace gml
