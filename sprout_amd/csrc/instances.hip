// Instanced meshes under a top-level tree (sprout's prop BVH):
//
//   BVH_wrapper::intersect / intersect_p   core/scene/prop/prop_bvh_wrapper.cpp:33-106, 183-247
//   Prop::intersect / intersect_p          core/scene/prop/prop.cpp:68-126 (prop AABB test first)
//   Mesh::intersect / intersect_p          core/scene/shape/triangle/triangle_mesh.cpp:366-446
//                                          (ray to object space, then the mesh's tree)
//   scene::bvh::Builder over prop AABBs    core/scene/bvh/scene_bvh_builder.cpp:14-107
//   prop AABB = shape AABB transformed     core/scene/scene.cpp:538-548, base/math/aabb.inl:263-280
//   transform_point / transform_vector     base/math/matrix4x4.inl:319-342
//
// The top-level tree is the same SBVH the mesh builder makes (both run
// Builder_base::split over reference boxes and serialize depth-first); it is
// built by the GPU builder over one degenerate triangle (lo, hi, lo) per
// instance, whose box is exactly the instance's world box.
//
// Traversal follows the reference loops: the top level walks the prop tree
// with the world ray; at a leaf each instance in slot order is box-tested
// with the current max_t, then the ray goes to object space (the compiled
// reference evaluates transform_point as fma(z, r2, fma(x, r0, y r1)) + r3)
// and walks the mesh tree on the same stack above a sentinel, as
// Tree::intersect does on the worker's node stack.

#include "instances.hpp"

#include "device_math.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace mi {

namespace {

using dev::F3;
using dev::Null;

constexpr uint32_t Stack_entries = 64;  // per lane in LDS; deeper scenes are refused (MI_ERR_STACK)
constexpr uint32_t Block         = 128;

struct Iray {
    F3       o;
    F3       d;
    F3       inv;
    float    min_t;
    float    max_t;
    uint32_t sign_bits;
};

__device__ __forceinline__ void set_inverse(Iray& r) {
    r.inv       = {dev::reciprocal_lane(r.d.x), dev::reciprocal_lane(r.d.y), dev::reciprocal_lane(r.d.z)};
    r.sign_bits = (r.inv.x < 0.f ? 1u : 0u) | (r.inv.y < 0.f ? 2u : 0u) | (r.inv.z < 0.f ? 4u : 0u);
}

using dev::xform_point;
using dev::xform_vector;

__device__ __forceinline__ uint32_t node_num(uint32_t meta) {
    return (meta >> 8) & 0xFFu;
}

__device__ __forceinline__ uint32_t node_axis(uint32_t meta) {
    return meta & 0xFFu;
}

struct Instance_stack {
    uint32_t* s;  // Stack_entries * blockDim.x words, entry-major
    uint32_t  sp;

    __device__ void push(uint32_t v) {
        if (sp < Stack_entries) {
            s[sp * blockDim.x + threadIdx.x] = v;
        }
        ++sp;
    }

    __device__ uint32_t pop() {
        --sp;
        return sp < Stack_entries ? s[sp * blockDim.x + threadIdx.x] : Null;
    }
};

// Tree::intersect / intersect_p of one mesh with an object-space ray, on the
// shared stack above a sentinel.  Closest: shrinks r.max_t on every hit.
template <bool Any>
__device__ bool mesh_trace(float4 const* __restrict__ nodes, float4 const* __restrict__ tris,
                           Iray& r, Instance_stack& stack, uint32_t& index, uint32_t& slot, float& u,
                           float& v) {
    bool hit = false;
    stack.push(Null);
    uint32_t n = 0;
    while (Null != n) {
        float4 const lo = nodes[2 * size_t(n) + 0];
        float4 const hi = nodes[2 * size_t(n) + 1];
        if (dev::slab(lo, hi, r.o, r.inv, r.min_t, r.max_t)) {
            uint32_t const data = __float_as_uint(lo.w);
            uint32_t const meta = __float_as_uint(hi.w);
            uint32_t const num  = node_num(meta);
            if (0 == num) {
                if (0 == ((r.sign_bits >> node_axis(meta)) & 1u)) {
                    stack.push(data + 1);
                    n = data;
                } else {
                    stack.push(data);
                    n = data + 1;
                }
                continue;
            }
            for (uint32_t i = data, end = data + num; i < end; ++i) {
                float4 const ta = tris[3 * size_t(i) + 0];
                float4 const t1 = tris[3 * size_t(i) + 1];
                float4 const t2 = tris[3 * size_t(i) + 2];
                float        t, tu, tv;
                if (dev::triangle({ta.x, ta.y, ta.z}, {t1.x, t1.y, t1.z}, {t2.x, t2.y, t2.z}, r.o,
                                  r.d, r.min_t, r.max_t, t, tu, tv)) {
                    hit = true;
                    if (Any) {
                        // unwind to the sentinel
                        while (Null != stack.pop()) {
                        }
                        return true;
                    }
                    r.max_t = t;
                    u       = tu;
                    v       = tv;
                    index   = __float_as_uint(ta.w);
                    slot    = i;
                }
            }
        }
        n = stack.pop();
    }
    return hit;
}

template <bool Any>
__global__ void __launch_bounds__(Block) trace_instances(Tlas_view t, mi_ray const* rays,
                                                         uint32_t num_rays,
                                                         uint32_t const* num_rays_dev, void* out) {
    extern __shared__ uint32_t lds[];
    Instance_stack stack;
    stack.s = lds;

    if (num_rays_dev) {
        num_rays = min(num_rays, *num_rays_dev);
    }
    uint32_t const stride = gridDim.x * blockDim.x;
    for (uint32_t ri = blockIdx.x * blockDim.x + threadIdx.x; ri < num_rays; ri += stride) {
        float4 const* rp = reinterpret_cast<float4 const*>(rays) + 2 * size_t(ri);
        float4 const  ro = rp[0];
        float4 const  rd = rp[1];
        Iray          w;
        w.o     = {ro.x, ro.y, ro.z};
        w.min_t = ro.w;
        w.d     = {rd.x, rd.y, rd.z};
        w.max_t = rd.w;
        set_inverse(w);

        uint32_t instance = Null;
        uint32_t index    = Null;
        uint32_t slot     = Null;
        float    u = 0.f, v = 0.f;
        bool     any = false;

        stack.sp = 0;
        stack.push(Null);
        uint32_t n = t.num_nodes ? 0u : Null;
        while (Null != n) {
            float4 const lo = t.nodes[2 * size_t(n) + 0];
            float4 const hi = t.nodes[2 * size_t(n) + 1];
            if (dev::slab(lo, hi, w.o, w.inv, w.min_t, w.max_t)) {
                uint32_t const data = __float_as_uint(lo.w);
                uint32_t const meta = __float_as_uint(hi.w);
                uint32_t const num  = node_num(meta);
                if (0 == num) {
                    if (0 == ((w.sign_bits >> node_axis(meta)) & 1u)) {
                        stack.push(data + 1);
                        n = data;
                    } else {
                        stack.push(data);
                        n = data + 1;
                    }
                    continue;
                }
                for (uint32_t i = data, end = data + num; i < end; ++i) {
                    uint32_t const        p  = __float_as_uint(t.slots[3 * size_t(i)].w);
                    Instance_record const ir = t.instances[p];
                    float4 const          blo = make_float4(ir.lo[0], ir.lo[1], ir.lo[2], 0.f);
                    float4 const          bhi = make_float4(ir.hi[0], ir.hi[1], ir.hi[2], 0.f);
                    // Prop::intersect: the prop's world box first
                    if (!dev::slab(blo, bhi, w.o, w.inv, w.min_t, w.max_t)) {
                        continue;
                    }
                    Iray o;
                    o.o     = xform_point(ir.world_to_object, w.o);
                    o.d     = xform_vector(ir.world_to_object, w.d);
                    o.min_t = w.min_t;
                    o.max_t = w.max_t;
                    set_inverse(o);
                    Mesh_view const mv = t.meshes[ir.mesh];
                    uint32_t        idx = Null, hslot = Null;
                    float           hu = 0.f, hv = 0.f;
                    if (mesh_trace<Any>(mv.nodes, mv.tris, o, stack, idx, hslot, hu, hv)) {
                        if (Any) {
                            any = true;
                            break;
                        }
                        w.max_t  = o.max_t;
                        instance = p;
                        index    = idx;
                        slot     = hslot;
                        u        = hu;
                        v        = hv;
                    }
                }
                if (Any && any) {
                    break;
                }
            }
            n = stack.pop();
        }

        if (Any) {
            static_cast<uint32_t*>(out)[ri] = any ? 1u : 0u;
        } else {
            mi_instance_hit h;
            h.t        = w.max_t;
            h.u        = Null != index ? u : 0.f;
            h.v        = Null != index ? v : 0.f;
            h.triangle = index;
            h.instance = instance;
            h.slot     = slot;
            h.pad[0] = h.pad[1] = 0;
            static_cast<mi_instance_hit*>(out)[ri] = h;
        }
    }
}

// AABB::transform (aabb.inl:263-280): per axis the products of the box's
// bounds with the matrix row, min / max, then (x + y) + (z + translation)
void transform_box(float const lo[3], float const hi[3], float const m[16], float out_lo[3],
                   float out_hi[3]) {
    for (int c = 0; c < 3; ++c) {
        float const xa = lo[0] * m[0 + c], xb = hi[0] * m[0 + c];
        float const ya = lo[1] * m[4 + c], yb = hi[1] * m[4 + c];
        float const za = lo[2] * m[8 + c], zb = hi[2] * m[8 + c];
        float const w  = m[12 + c];
        out_lo[c]      = (std::min(xa, xb) + std::min(ya, yb)) + (std::min(za, zb) + w);
        out_hi[c]      = (std::max(xa, xb) + std::max(ya, yb)) + (std::max(za, zb) + w);
    }
}

}  // namespace

void Tlas::release() {
    tree.release();
    (void)hipFree(instances);
    (void)hipFree(meshes);
    instances = nullptr;
    meshes    = nullptr;
    records.clear();
    mesh_handles.clear();
    num_instances = 0;
    max_stack     = 0;
}

int32_t create_tlas(DeviceBvh const* const* meshes, uint32_t num_meshes, uint32_t const* mesh_of,
                    float const* world_to_object, float const* object_to_world,
                    uint32_t num_instances, mi_bvh_settings const& settings, hipStream_t stream,
                    Tlas& out) {
    out.release();
    if (0 == num_instances) {
        return fail(MI_ERR_ARGS, "instances: none given");
    }
    std::vector<float>    pos(6 * size_t(num_instances));
    std::vector<uint32_t> idx(3 * size_t(num_instances));
    out.records.resize(num_instances);
    uint32_t deepest = 0;
    for (uint32_t i = 0; i < num_instances; ++i) {
        if (mesh_of[i] >= num_meshes) {
            return fail(MI_ERR_ARGS, "instance %u: mesh %u out of range", i, mesh_of[i]);
        }
        DeviceBvh const& m = *meshes[mesh_of[i]];
        Instance_record& r = out.records[i];
        std::memcpy(r.world_to_object, world_to_object + 16 * size_t(i), 16 * sizeof(float));
        for (int k = 0; k < 16; ++k) {
            if (!std::isfinite(r.world_to_object[k]) || !std::isfinite(object_to_world[16 * size_t(i) + k])) {
                return fail(MI_ERR_ARGS, "instance %u: matrix not finite", i);
            }
        }
        transform_box(m.aabb_min, m.aabb_max, object_to_world + 16 * size_t(i), r.lo, r.hi);
        r.mesh = mesh_of[i];
        for (int k = 0; k < 3; ++k) {
            pos[6 * size_t(i) + k]     = r.lo[k];
            pos[6 * size_t(i) + 3 + k] = r.hi[k];
        }
        idx[3 * size_t(i) + 0] = 2 * i;
        idx[3 * size_t(i) + 1] = 2 * i + 1;
        idx[3 * size_t(i) + 2] = 2 * i;
        deepest = std::max(deepest, m.max_depth);
    }

    float*    d_pos = nullptr;
    uint32_t* d_idx = nullptr;
    hipError_t e    = hipMalloc(&d_pos, pos.size() * sizeof(float));
    if (!e) e = hipMalloc(&d_idx, idx.size() * sizeof(uint32_t));
    if (!e) e = hipMemcpy(d_pos, pos.data(), pos.size() * sizeof(float), hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(d_idx, idx.data(), idx.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e) {
        (void)hipFree(d_pos);
        (void)hipFree(d_idx);
        return fail(MI_ERR_HIP, "instances: %s", hipGetErrorString(e));
    }
    int32_t st = build_bvh(d_pos, 2 * num_instances, 3, d_idx, num_instances, settings, stream, out.tree);
    (void)hipFree(d_pos);
    (void)hipFree(d_idx);
    if (st) {
        out.release();
        return st;
    }

    // a full stack holds the top level's pending nodes, a sentinel and the
    // mesh's (the reference's Node_stack has 128 entries)
    out.max_stack = out.tree.max_depth + deepest + 3;
    if (out.max_stack > Stack_entries) {
        out.release();
        return fail(MI_ERR_STACK, "instances: top depth + mesh depth exceed %u stack entries",
                    Stack_entries);
    }

    std::vector<Mesh_view> views(num_meshes);
    for (uint32_t k = 0; k < num_meshes; ++k) {
        views[k].nodes = meshes[k]->nodes;
        views[k].tris  = meshes[k]->tris;
    }
    e = hipMalloc(&out.instances, num_instances * sizeof(Instance_record));
    if (!e) e = hipMalloc(&out.meshes, num_meshes * sizeof(Mesh_view));
    if (!e) e = hipMemcpy(out.instances, out.records.data(), num_instances * sizeof(Instance_record),
                          hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(out.meshes, views.data(), num_meshes * sizeof(Mesh_view), hipMemcpyHostToDevice);
    if (e) {
        out.release();
        return fail(MI_ERR_HIP, "instances: %s", hipGetErrorString(e));
    }
    out.num_instances = num_instances;
    out.num_meshes    = num_meshes;
    return MI_OK;
}

int32_t trace_instances_launch(Tlas const& t, uint32_t num_cus, mi_ray const* rays, uint32_t n,
                               uint32_t const* n_dev, void* out, bool any_hit, hipStream_t stream) {
    if (0 == n) {
        return MI_OK;
    }
    Tlas_view const v      = tlas_view(t);
    uint64_t const  wanted = (uint64_t(n) + Block - 1) / Block;
    uint32_t const blocks = uint32_t(std::min<uint64_t>(wanted, uint64_t(num_cus) * 8));
    size_t const   lds    = size_t(Stack_entries) * Block * sizeof(uint32_t);
    if (any_hit) {
        hipLaunchKernelGGL(trace_instances<true>, dim3(blocks), dim3(Block), lds, stream, v, rays, n,
                           n_dev, out);
    } else {
        hipLaunchKernelGGL(trace_instances<false>, dim3(blocks), dim3(Block), lds, stream, v, rays, n,
                           n_dev, out);
    }
    if (hipError_t const e = hipGetLastError(); e) {
        return fail(MI_ERR_HIP, "instance trace: %s", hipGetErrorString(e));
    }
    return MI_OK;
}

}  // namespace mi
