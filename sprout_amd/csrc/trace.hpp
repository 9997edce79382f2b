// Device-side tree and traversal launch interface (host code).
#pragma once

#include "common.hpp"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace mi {

// A tree resident in HBM; see trace.hip for the layout.
struct DeviceBvh {
    void*     nodes_alloc = nullptr;  // 64-byte aligned; nodes start 32 bytes in
    float4*   nodes       = nullptr;  // 2 float4 per node
    float4*   tris        = nullptr;  // 3 float4 per tree slot
    uint32_t* slots       = nullptr;  // input triangle id per tree slot

    uint32_t num_nodes     = 0;
    uint32_t num_slots     = 0;
    uint32_t num_triangles = 0;
    uint32_t max_depth     = 0;

    float  aabb_min[3] = {0.f, 0.f, 0.f};
    float  aabb_max[3] = {0.f, 0.f, 0.f};
    double build_ms    = 0.0;

    // Compressed sibling pairs of kernel 4, made on first use: per pair of
    // children (a, a + 1) 32 bytes, the two boxes as 16-bit grid coordinates
    // (lower corners rounded down, upper ones up, so every box holds the
    // original) and the two packed node words; the grid spans the root box.
    mutable uint4* qpairs     = nullptr;
    mutable float  qorigin[3] = {0.f, 0.f, 0.f};
    mutable float  qscale[3]  = {0.f, 0.f, 0.f};
    mutable bool   qfailed    = false;  // a child pair starts at an even index: no compressed copy

    void release();
};

// The refill traversals claim rays from Claim_parts counter words, each on
// its own 256-byte line: a ray counter buffer holds Counter_words words.
inline constexpr uint32_t Claim_parts   = 32;
inline constexpr uint32_t Claim_stride  = 64;
inline constexpr uint32_t Counter_words = Claim_parts * Claim_stride;

// Persistent-grid launch state shared by all traversals of a context.
struct TraceResources {
    uint32_t  block_size    = 256;
    uint32_t  blocks_per_cu = 8;
    uint32_t  num_cus       = 256;
    uint32_t  kernel        = 3;     // 0 refill, 1 static grid-stride, 2 refill + while-while,
                                     // 3 as 2 with two-word stack entries (small enough trees),
                                     // 4 as 3 over compressed 32-byte sibling pairs
    uint32_t  refill_lanes  = 8;     // hand out rays once this many lanes of a wave are idle
    uint32_t  claim_batch   = 64;    // rays a wave claims from the global counter at a time
    uint32_t  lds_entries   = 10;    // stack entries kept in LDS (0: up to 32); deeper ones spill
    uint32_t* counter       = nullptr;  // ray claim counters of trace_refill (Counter_words)
    uint32_t* spill         = nullptr;
    size_t    spill_words   = 0;
    // kernels 3 and 4 with a host ray count: trace the rays in the order of
    // a key (direction octant, Morton code of the origin in the tree's box)
    // so that a wave's rays start close together; results land in the
    // callers' order, so they are the same
    uint32_t  sort_rays     = 0;
    uint32_t* sort_keys     = nullptr;  // 4 arrays of sort_capacity words: keys in / out, order in / out
    void*     sort_temp     = nullptr;
    size_t    sort_temp_bytes = 0;
    uint32_t  sort_capacity   = 0;

    uint32_t grid_blocks(uint32_t num_rays) const {
        uint64_t const wanted   = (uint64_t(num_rays) + block_size - 1) / block_size;
        uint64_t const resident = uint64_t(num_cus) * blocks_per_cu;
        uint64_t const b        = wanted < resident ? wanted : resident;
        return uint32_t(b > 0 ? b : 1);
    }

    int32_t reserve_spill(size_t words);

    void release();
};

// Allocates the device tree from host arrays in sprout's serialized layout
// (nodes: 32-byte scene::bvh::Node; slots: input triangle per tree slot) and
// a host mesh; computes depth and bounds.
int32_t upload_bvh(void const* nodes, uint32_t num_nodes, uint32_t const* slots,
                   uint32_t num_slots, float const* positions, uint32_t num_vertices,
                   uint32_t position_stride, uint32_t const* indices, uint32_t num_triangles,
                   DeviceBvh& out);

// Fills bvh.tris from device-resident slots and mesh.
int32_t fill_triangles(DeviceBvh& bvh, float const* positions, uint32_t position_stride,
                       uint32_t const* indices, hipStream_t stream);

int32_t launch_trace(DeviceBvh const& bvh, TraceResources& res, mi_ray const* rays,
                     uint32_t num_rays, void* out, uint32_t* counts, bool any_hit,
                     hipStream_t stream);

// Reserves what traversals of up to max_rays rays need (the stack spill
// area), so that later launches allocate nothing (as inside graph capture).
int32_t prepare_trace(DeviceBvh const& bvh, TraceResources& res, uint32_t max_rays);

// Same traversal over a ray count that lives in device memory (written by an
// earlier kernel on the stream); max_rays bounds it.  Capturable: it records
// kernels only.
int32_t launch_trace_device_count(DeviceBvh const& bvh, TraceResources& res, mi_ray const* rays,
                                  uint32_t const* num_rays, uint32_t max_rays, void* out,
                                  bool any_hit, hipStream_t stream);

// Sets up the fixed scratch of launch_trace_noalloc on `device`, which must
// be the current device (once per device; later calls return at once).
int32_t noalloc_scratch_setup(int device);

// Closest hit as launch_trace, allocating nothing: the ray counters and stack
// spill live in the library's fixed 32 MiB device array (the persistent grid
// shrinks to fit it).  Calls that overlap on one device race on that array.
int32_t launch_trace_noalloc(DeviceBvh const& bvh, TraceResources& res, int device,
                             mi_ray const* rays, uint32_t num_rays, mi_hit* hits,
                             hipStream_t stream);

}  // namespace mi
