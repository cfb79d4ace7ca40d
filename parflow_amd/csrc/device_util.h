// device_util.h -- small device helpers shared by the HIP sources.
#ifndef PFHIP_DEVICE_UTIL_H
#define PFHIP_DEVICE_UTIL_H

#include <hip/hip_runtime.h>

// Workgroup barrier that orders LDS only: waits for this wave's LDS
// operations, not for its outstanding global loads (prefetches stay in
// flight across it).  The "memory" clobber keeps the compiler from moving
// memory accesses across it.  Global memory written before it is NOT made
// visible to other waves by it.
__device__ __forceinline__ void lds_barrier()
{
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

#endif
