// Grid-wide synchronisation for kernels whose workgroups cooperate on one
// system (a grid of at most one workgroup per CU per system, so that every
// workgroup is resident).  Per system, a header of kCoopHeaderWords 64-bit
// words in device memory: word 0 counts barrier arrivals, word 8 is the
// abort flag.  Exchange data around a barrier moves with agent-scope atomic
// loads and stores; the counter is released and acquired at agent scope.  A
// wait longer than the spin limit sets the abort flag, and every later
// barrier of that system returns false at once, so the kernel always ends.
#pragma once
#include "ccab_internal.h"

namespace {

constexpr int kCoopHeaderWords = 32;
// polls (each an agent-scope load plus a ~128-cycle sleep) a barrier may
// wait, a few seconds, before it gives up
constexpr unsigned kCoopSpinLimit = 1u << 21;

__global__ void coop_reset_kernel(uint64_t* xbuf, int64_t sys_words, int64_t batch) {
    const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    xbuf[b * sys_words] = 0;       // barrier counter
    xbuf[b * sys_words + 8] = 0;   // abort flag
    xbuf[b * sys_words + 16] = 0;  // leave counter (kernels that clear their own header)
}

// Grid barrier of one system: waits until the counter reaches target (all G
// workgroups have arrived at this barrier).  Returns false, uniformly in the
// workgroup, when the system's abort flag is set (by this or another
// workgroup's timed-out wait).
__device__ __forceinline__ bool coop_barrier(uint64_t* hdr, uint64_t target, unsigned spin_limit,
                                             int* s_ok) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(hdr, uint64_t(1), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        unsigned spins = 0;
        int ok = 1;
        while (__hip_atomic_load(hdr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(hdr + 8, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                ok = 0;
                break;
            }
            if (++spins > spin_limit) {
                __hip_atomic_store(hdr + 8, uint64_t(1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *s_ok = ok;
    }
    __syncthreads();
    return *s_ok != 0;
}

}  // namespace
