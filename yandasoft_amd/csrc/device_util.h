// Device helpers shared by the kernel files of libygrid.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ygrid_internal.h"

namespace ygrid {

// FP32 -> FP16 or BF16 bits, round to nearest even (finite values; BF16 by the carry into the
// upper half, FP16 by the hardware conversion), and back.
__device__ __forceinline__ uint16_t round16(float x, int prec) {
  if (prec == kPrecF16) return __builtin_bit_cast(uint16_t, (_Float16)x);
  const uint32_t u = __builtin_bit_cast(uint32_t, x);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float widen16(uint32_t h, int prec) {
  if (prec == kPrecF16) return (float)__builtin_bit_cast(_Float16, (uint16_t)h);
  return __builtin_bit_cast(float, h << 16);
}

// Work-item queue of a persistent launch: nitems items split into kItemQueues contiguous ranges,
// one counter each (kQueueStride ints apart); a workgroup starts on queue q (its XCD's, by
// blockIdx) and moves on to the next queue when its own is drained.  Returns the item (uniform
// over the workgroup), or -1 once every queue is drained.  A queue seen drained by a plain load
// is passed without an atomic (counters only grow).  Contains two workgroup barriers.
__device__ __forceinline__ int next_item(int* __restrict__ counter, int nitems, int& q, int& left, int* s_item) {
  for (;;) {
    const int lo = (int)((long long)nitems * q / kItemQueues);
    const int hi = (int)((long long)nitems * (q + 1) / kItemQueues);
    if (threadIdx.x == 0) {
      int* cq = counter + q * kQueueStride;
      const int seen = __hip_atomic_load(cq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      *s_item = lo + seen >= hi ? hi : lo + atomicAdd(cq, 1);
    }
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(*s_item);  // uniform: item data via scalar loads
    __syncthreads();
    if (item < hi) return item;
    if (--left == 0) return -1;
    q = q + 1 == kItemQueues ? 0 : q + 1;
  }
}

}  // namespace ygrid
