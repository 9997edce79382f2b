// Device tree storage: allocation, upload of host-built trees and the
// per-slot triangle records the traversal reads.

#include "common.hpp"
#include "trace.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

namespace mi {

namespace {

// One 48-byte record per tree slot: {a.xyz, input triangle}{b - a}{c - a}.
// Slot order is the tree's leaf order (triangle_bvh_builder_sah.cpp:152-160).
__global__ void triangle_records(uint32_t const* __restrict__ slots, uint32_t num_slots,
                                 float const* __restrict__ positions, uint32_t stride,
                                 uint32_t const* __restrict__ indices, float4* __restrict__ tris) {
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num_slots) {
        return;
    }

    uint32_t const t  = slots[i];
    float const*   pa = positions + size_t(indices[3 * size_t(t) + 0]) * stride;
    float const*   pb = positions + size_t(indices[3 * size_t(t) + 1]) * stride;
    float const*   pc = positions + size_t(indices[3 * size_t(t) + 2]) * stride;

    tris[3 * size_t(i) + 0] = make_float4(pa[0], pa[1], pa[2], __uint_as_float(t));
    tris[3 * size_t(i) + 1] = make_float4(pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2], 0.f);
    tris[3 * size_t(i) + 2] = make_float4(pc[0] - pa[0], pc[1] - pa[1], pc[2] - pa[2], 0.f);
}

struct HostNode {
    float    lo[3];
    uint32_t data;
    float    hi[3];
    uint8_t  axis;
    uint8_t  num;
    uint8_t  pad[2];
};
static_assert(sizeof(HostNode) == 32);

}  // namespace

void DeviceBvh::release() {
    (void)hipFree(qpairs);
    qpairs = nullptr;
    (void)hipFree(nodes_alloc);
    (void)hipFree(tris);
    (void)hipFree(slots);
    nodes_alloc = nullptr;
    nodes       = nullptr;
    tris        = nullptr;
    slots       = nullptr;
}

int32_t TraceResources::reserve_spill(size_t words) {
    if (words <= spill_words) {
        return MI_OK;
    }
    (void)hipFree(spill);
    spill       = nullptr;
    spill_words = 0;
    if (hipError_t const e = hipMalloc(&spill, words * sizeof(uint32_t)); e != hipSuccess) {
        return fail(MI_ERR_HIP, "spill stack allocation of %zu words: %s", words,
                    hipGetErrorString(e));
    }
    spill_words = words;
    return MI_OK;
}

void TraceResources::release() {
    (void)hipFree(spill);
    (void)hipFree(counter);
    (void)hipFree(sort_keys);
    (void)hipFree(sort_temp);
    spill           = nullptr;
    counter         = nullptr;
    spill_words     = 0;
    sort_keys       = nullptr;
    sort_temp       = nullptr;
    sort_temp_bytes = 0;
    sort_capacity   = 0;
}

int32_t fill_triangles(DeviceBvh& bvh, float const* positions, uint32_t position_stride,
                       uint32_t const* indices, hipStream_t stream) {
    if (0 == bvh.num_slots) {
        return MI_OK;
    }
    uint32_t const block = 256;
    uint32_t const grid  = (bvh.num_slots + block - 1) / block;
    hipLaunchKernelGGL(triangle_records, dim3(grid), dim3(block), 0, stream, bvh.slots,
                       bvh.num_slots, positions, position_stride, indices, bvh.tris);
    if (hipError_t const e = hipGetLastError(); e != hipSuccess) {
        return fail(MI_ERR_HIP, "triangle_records launch: %s", hipGetErrorString(e));
    }
    return MI_OK;
}

int32_t upload_bvh(void const* nodes, uint32_t num_nodes, uint32_t const* slots,
                   uint32_t num_slots, float const* positions, uint32_t num_vertices,
                   uint32_t position_stride, uint32_t const* indices, uint32_t num_triangles,
                   DeviceBvh& out) {
    HostNode const* hn = static_cast<HostNode const*>(nodes);

    // Structural checks: every index stays inside its array and children
    // follow their parent, so a traversal cannot leave the tree or loop.  A
    // node's depth is the longest path to it (parents precede children), which
    // bounds the stack a traversal needs even for a shared child.
    std::vector<uint32_t> depth(num_nodes, 0);
    uint32_t              max_depth = 0;
    for (uint32_t i = 0; i < num_nodes; ++i) {
        HostNode const& n = hn[i];
        if (0 == n.num) {
            if (n.axis > 2 || n.data <= i || uint64_t(n.data) + 1 >= num_nodes) {
                return fail(MI_ERR_ARGS, "node %u: bad children %u (axis %u)", i, n.data, n.axis);
            }
            depth[n.data]     = std::max(depth[n.data], depth[i] + 1);
            depth[n.data + 1] = std::max(depth[n.data + 1], depth[i] + 1);
            max_depth         = std::max(max_depth, depth[i] + 1);
        } else if (uint64_t(n.data) + n.num > num_slots) {
            return fail(MI_ERR_ARGS, "leaf %u: slots [%u, %u) beyond %u", i, n.data,
                        n.data + n.num, num_slots);
        }
    }
    for (uint32_t i = 0; i < num_slots; ++i) {
        if (slots[i] >= num_triangles) {
            return fail(MI_ERR_ARGS, "slot %u: triangle %u beyond %u", i, slots[i], num_triangles);
        }
    }
    for (uint64_t i = 0; i < 3ull * num_triangles; ++i) {
        if (indices[i] >= num_vertices) {
            return fail(MI_ERR_ARGS, "index %llu: vertex %u beyond %u", (unsigned long long)i,
                        indices[i], num_vertices);
        }
    }

    DeviceBvh bvh;
    bvh.num_nodes     = num_nodes;
    bvh.num_slots     = num_slots;
    bvh.num_triangles = num_triangles;
    bvh.max_depth     = max_depth;
    for (int k = 0; k < 3; ++k) {
        bvh.aabb_min[k] = hn[0].lo[k];
        bvh.aabb_max[k] = hn[0].hi[k];
    }

    float*    d_pos = nullptr;
    uint32_t* d_idx = nullptr;

    size_t const pos_bytes = size_t(num_vertices) * position_stride * sizeof(float);
    size_t const idx_bytes = size_t(num_triangles) * 3 * sizeof(uint32_t);

    hipError_t e = hipMalloc(&bvh.nodes_alloc, size_t(num_nodes + 1) * 32);
    if (e == hipSuccess) {
        e = hipMalloc(&bvh.tris, size_t(num_slots ? num_slots : 1) * 48);
    }
    if (e == hipSuccess) {
        e = hipMalloc(&bvh.slots, size_t(num_slots ? num_slots : 1) * 4);
    }
    if (e == hipSuccess) {
        e = hipMalloc(&d_pos, pos_bytes ? pos_bytes : 4);
    }
    if (e == hipSuccess) {
        e = hipMalloc(&d_idx, idx_bytes ? idx_bytes : 4);
    }
    if (e == hipSuccess) {
        bvh.nodes = reinterpret_cast<float4*>(static_cast<char*>(bvh.nodes_alloc) + 32);
        e         = hipMemcpy(bvh.nodes, nodes, size_t(num_nodes) * 32, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && num_slots) {
        e = hipMemcpy(bvh.slots, slots, size_t(num_slots) * 4, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && pos_bytes) {
        e = hipMemcpy(d_pos, positions, pos_bytes, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && idx_bytes) {
        e = hipMemcpy(d_idx, indices, idx_bytes, hipMemcpyHostToDevice);
    }

    int32_t st = MI_OK;
    if (e != hipSuccess) {
        st = fail(MI_ERR_HIP, "bvh upload: %s", hipGetErrorString(e));
    } else {
        st = fill_triangles(bvh, d_pos, position_stride, d_idx, nullptr);
        if (st == MI_OK) {
            if (hipError_t const s = hipDeviceSynchronize(); s != hipSuccess) {
                st = fail(MI_ERR_HIP, "bvh upload: %s", hipGetErrorString(s));
            }
        }
    }

    (void)hipFree(d_pos);
    (void)hipFree(d_idx);

    if (st != MI_OK) {
        bvh.release();
        return st;
    }

    out = bvh;
    return MI_OK;
}

}  // namespace mi
