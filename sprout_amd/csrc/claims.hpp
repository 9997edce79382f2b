// Ray claims of the persistent refill traversals (device code).
#pragma once

#include "trace.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mi {

// The ray claims of the refill kernels.  The rays are split into
// Claim_parts equal parts, each with its own counter word on its own 256-byte
// line (Claim_stride words apart).  A wave starts on part (wave mod
// Claim_parts) and moves to the next part once its part is used up.
// Atomics on one address serialize (about 12 ns each), so spreading the
// claims over the parts keeps them off the critical path.
struct Claims {
    uint32_t part;   // wave-uniform: the part claimed from next
    uint32_t tried;  // parts found used up (they stay used up)

    __device__ void init() {
        part  = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) % Claim_parts;
        tried = 0;
    }

    // Claims up to batch rays [begin, end) for the wave; the leader lane
    // issues the atomic.  False once every part is used up.
    __device__ bool next(uint32_t* counters, uint32_t num_rays, uint32_t batch, uint32_t lane,
                         uint32_t leader, uint32_t& begin, uint32_t& end) {
        uint64_t const size = (uint64_t(num_rays) + Claim_parts - 1) / Claim_parts;
        while (tried < Claim_parts) {
            uint64_t const  lo_part = uint64_t(part) * size;
            uint64_t const  lo      = lo_part < num_rays ? lo_part : num_rays;
            uint64_t const  hi      = lo + size < num_rays ? lo + size : num_rays;
            uint32_t* const c       = counters + part * Claim_stride;
            uint32_t        taken   = 0;
            if (lane == leader) {
                // a part seen used up costs no atomic
                taken = __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (lo + taken < hi) {
                    taken = atomicAdd(c, batch);
                }
            }
            taken = __shfl(taken, leader);
            if (lo + taken < hi) {
                begin = uint32_t(lo + taken);
                end   = uint32_t(lo + taken + batch < hi ? lo + taken + batch : hi);
                return true;
            }
            part = part + 1 == Claim_parts ? 0u : part + 1;
            ++tried;
        }
        return false;
    }
};

}  // namespace mi
