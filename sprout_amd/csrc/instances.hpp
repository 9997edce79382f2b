// Instanced meshes under a top-level tree (host interface); see instances.hip.
#pragma once

#include "bvh_build.hpp"
#include "common.hpp"
#include "device_math.hpp"
#include "trace.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace mi {

// One instance on the device: world-to-object rows (row-vector 4x4, as
// sprout's Composed_transformation::world_to_object), world box, mesh.
struct Instance_record {
    float    world_to_object[16];
    float    lo[3];
    float    hi[3];
    uint32_t mesh;
    uint32_t pad;
};

namespace dev {

// transform_point / transform_vector (base/math/matrix4x4.inl:319-342) with
// the rounding the reference build performs: fma(z, r2, fma(x, r0, y r1))
// (+ r3), rows r0..r3 of a row-vector 4x4
__device__ __forceinline__ F3 xform_point(float const* m, F3 v) {
    F3 r;
    r.x = __builtin_fmaf(v.z, m[8], __builtin_fmaf(v.x, m[0], v.y * m[4])) + m[12];
    r.y = __builtin_fmaf(v.z, m[9], __builtin_fmaf(v.x, m[1], v.y * m[5])) + m[13];
    r.z = __builtin_fmaf(v.z, m[10], __builtin_fmaf(v.x, m[2], v.y * m[6])) + m[14];
    return r;
}

__device__ __forceinline__ F3 xform_vector(float const* m, F3 v) {
    F3 r;
    r.x = __builtin_fmaf(v.z, m[8], __builtin_fmaf(v.x, m[0], v.y * m[4]));
    r.y = __builtin_fmaf(v.z, m[9], __builtin_fmaf(v.x, m[1], v.y * m[5]));
    r.z = __builtin_fmaf(v.z, m[10], __builtin_fmaf(v.x, m[2], v.y * m[6]));
    return r;
}

}  // namespace dev

struct Mesh_view {
    float4 const* nodes;
    float4 const* tris;
};

struct Tlas_view {
    float4 const*          nodes;
    float4 const*          slots;  // 3 float4 per top-level slot; .w of the first = instance
    uint32_t               num_nodes;
    Instance_record const* instances;
    Mesh_view const*       meshes;
};

struct Tlas {
    DeviceBvh                    tree;  // over one box per instance
    Instance_record*             instances = nullptr;
    Mesh_view*                   meshes    = nullptr;
    std::vector<Instance_record> records;       // host copy
    std::vector<uint32_t>        mesh_handles;  // the BVH handles the instances use
    uint32_t                     num_instances = 0;
    uint32_t                     num_meshes    = 0;
    uint32_t                     max_stack     = 0;
    uint32_t*                    spill         = nullptr;  // stack area of the pair traversal
    uint32_t*                    claims        = nullptr;  // its ray claim counters (Counter_words)
    uint32_t                     spill_total   = 0;        // its threads (0: not used)

    void release();
};

inline Tlas_view tlas_view(Tlas const& t) {
    Tlas_view v;
    v.nodes     = t.tree.nodes;
    v.slots     = t.tree.tris;
    v.num_nodes = t.tree.num_nodes;
    v.instances = t.instances;
    v.meshes    = t.meshes;
    return v;
}

// meshes[k]: device trees; instance i uses meshes[mesh_of[i]] with the
// row-vector 4x4 matrices world_to_object / object_to_world (16 floats each,
// host).  Builds the top-level tree on the device.
int32_t create_tlas(DeviceBvh const* const* meshes, uint32_t num_meshes, uint32_t const* mesh_of,
                    float const* world_to_object, float const* object_to_world,
                    uint32_t num_instances, mi_bvh_settings const& settings, hipStream_t stream,
                    Tlas& out);

// Closest (out: mi_instance_hit per ray, pad[0] = the mesh tree's slot) or
// any hit (out: uint32 per ray).  n_dev (optional): a ray count in device
// memory, at most n (capturable).
int32_t trace_instances_launch(Tlas const& t, uint32_t num_cus, mi_ray const* rays, uint32_t n,
                               uint32_t const* n_dev, void* out, bool any_hit, hipStream_t stream);

}  // namespace mi
