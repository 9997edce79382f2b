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

#include "claims.hpp"
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

// ---- two-word stack traversal ----
//
// The same visiting order with 64-byte sibling-pair fetches: at an inner
// node both children are box-tested, the far one (when both are hit) is
// pushed as {its node word, its box entry distance} and the near one taken.
// A popped child is skipped when its entry distance exceeds the current
// max_t, which is exactly the reference's box test of it at that max_t
// (slab_entry).  A node word packs data (< 2^22), axis and count, so popped
// children need no node fetch.  The mesh walk runs above a sentinel word on
// the same stack.  The first Lds_entries entries of a lane live in LDS, the
// rest in a per-thread global area.

constexpr uint32_t Lds_entries   = 12;
constexpr uint32_t Block2        = 64;
constexpr uint32_t Blocks_per_cu = 24;
constexpr uint32_t Data_bits     = 22;
constexpr uint32_t Sentinel_word = 0xFFFFFFFFu;  // no valid node word (its slot range ends past 2^22)

__device__ __forceinline__ uint32_t pack_node(uint32_t data, uint32_t meta) {
    return data | ((node_axis(meta) & 3u) << Data_bits) | (node_num(meta) << 24);
}

__device__ __forceinline__ void unpack_node(uint32_t e, uint32_t& data, uint32_t& meta) {
    data = e & ((1u << Data_bits) - 1u);
    meta = ((e >> Data_bits) & 3u) | ((e >> 24) << 8);
}

struct Pair_stack {
    uint32_t* lds;    // 2 * Lds_entries * Block2 words, entry-major
    uint32_t* spill;  // 2 * spill entries * total words, entry-major
    uint32_t  gid;
    uint32_t  total;
    uint32_t  sp;

    __device__ void push(uint32_t node, float tmin) {
        if (sp < Lds_entries) {
            lds[(2 * sp) * Block2 + threadIdx.x]     = node;
            lds[(2 * sp + 1) * Block2 + threadIdx.x] = __float_as_uint(tmin);
        } else {
            size_t const k                  = 2 * size_t(sp - Lds_entries);
            spill[k * total + gid]       = node;
            spill[(k + 1) * total + gid] = __float_as_uint(tmin);
        }
        ++sp;
    }

    __device__ void pop(uint32_t& node, float& tmin) {
        --sp;
        if (sp < Lds_entries) {
            node = lds[(2 * sp) * Block2 + threadIdx.x];
            tmin = __uint_as_float(lds[(2 * sp + 1) * Block2 + threadIdx.x]);
        } else {
            size_t const k = 2 * size_t(sp - Lds_entries);
            node           = spill[k * total + gid];
            tmin           = __uint_as_float(spill[(k + 1) * total + gid]);
        }
    }
};

// Pops until a pending child passes its box test; false at a sentinel (which
// is consumed) or an empty stack.
__device__ __forceinline__ bool pop_pending(Pair_stack& stack, Iray const& r, uint32_t& data,
                                            uint32_t& meta) {
    while (stack.sp > 0) {
        uint32_t e;
        float    tmin;
        stack.pop(e, tmin);
        if (Sentinel_word == e) {
            return false;
        }
        if (r.max_t >= tmin) {
            unpack_node(e, data, meta);
            return true;
        }
    }
    return false;
}

// One inner node of either level (node data = first of the sibling pair).
__device__ __forceinline__ bool pair_step(float4 const* __restrict__ nodes, Pair_stack& stack,
                                          Iray const& r, uint32_t& data, uint32_t& meta) {
    uint32_t const a   = data;
    float4 const   alo = nodes[2 * size_t(a) + 0];
    float4 const   ahi = nodes[2 * size_t(a) + 1];
    float4 const   blo = nodes[2 * size_t(a) + 2];
    float4 const   bhi = nodes[2 * size_t(a) + 3];
    float          ta, tb;
    bool const     ha   = dev::slab_entry(alo, ahi, r.o, r.inv, r.min_t, r.max_t, ta);
    bool const     hb   = dev::slab_entry(blo, bhi, r.o, r.inv, r.min_t, r.max_t, tb);
    bool const     flip = (r.sign_bits >> node_axis(meta)) & 1u;
    bool const     hn   = flip ? hb : ha;
    bool const     hf   = flip ? ha : hb;
    if (hn | hf) {
        if (hn & hf) {
            stack.push(pack_node(__float_as_uint(flip ? alo.w : blo.w),
                                 __float_as_uint(flip ? ahi.w : bhi.w)),
                       flip ? ta : tb);
        }
        bool const take_b = hn ? flip : !flip;
        data              = __float_as_uint(take_b ? blo.w : alo.w);
        meta              = __float_as_uint(take_b ? bhi.w : ahi.w);
        return true;
    }
    return pop_pending(stack, r, data, meta);
}

// Tree::intersect / intersect_p of one mesh above a sentinel; returns with
// the sentinel popped.  Closest: shrinks r.max_t on every hit.
template <bool Any>
__device__ bool mesh_pairs(float4 const* __restrict__ nodes, float4 const* __restrict__ tris,
                           Iray& r, Pair_stack& stack, uint32_t& index, uint32_t& slot, float& u,
                           float& v) {
    stack.push(Sentinel_word, 0.f);
    bool         hit  = false;
    float4 const rlo  = nodes[0];
    float4 const rhi  = nodes[1];
    uint32_t     data = __float_as_uint(rlo.w);
    uint32_t     meta = __float_as_uint(rhi.w);
    bool         live = dev::slab(rlo, rhi, r.o, r.inv, r.min_t, r.max_t);
    if (!live) {
        uint32_t e;
        float    t;
        stack.pop(e, t);  // the sentinel
        return false;
    }
    while (live) {
        uint32_t const num = node_num(meta);
        if (0 == num) {
            live = pair_step(nodes, stack, r, data, meta);
            continue;
        }
        for (uint32_t i = data, end = data + num; i < end; ++i) {
            float4 const ta = tris[3 * size_t(i) + 0];
            float4 const t1 = tris[3 * size_t(i) + 1];
            float4 const t2 = tris[3 * size_t(i) + 2];
            float        t, tu, tv;
            if (dev::triangle({ta.x, ta.y, ta.z}, {t1.x, t1.y, t1.z}, {t2.x, t2.y, t2.z}, r.o, r.d,
                              r.min_t, r.max_t, t, tu, tv)) {
                hit = true;
                if (Any) {
                    // unwind to the sentinel
                    uint32_t e;
                    float    tm;
                    do {
                        stack.pop(e, tm);
                    } while (Sentinel_word != e);
                    return true;
                }
                r.max_t = t;
                u       = tu;
                v       = tv;
                index   = __float_as_uint(ta.w);
                slot    = i;
            }
        }
        live = pop_pending(stack, r, data, meta);
    }
    return hit;
}

// The two-level traversal with the two-word stack as a persistent kernel
// with per-lane ray refill: each lane walks its ray as a small state machine
// (a top-level node, the instances of a top-level leaf in slot order, or a
// node of an instance's mesh above a sentinel), one step per loop turn, and
// takes a new ray when its ray is done, so a wave's lanes stay busy while
// rays of very different lengths finish (one ray per lane for the whole walk
// ran 0.54 Grays/s on 64 instances; this runs 1.49).  Each ray visits nodes,
// instances and triangles in the reference's order; waves claim rays from
// spread counters (Claims).
constexpr uint32_t Refill_lanes = 8;   // refill once this many lanes of a wave are idle
constexpr uint32_t Claim_batch  = 64;  // rays a wave claims per atomic

constexpr uint32_t Phase_top  = 0;  // at a node of the top-level tree
constexpr uint32_t Phase_leaf = 1;  // box-testing the instances of a top-level leaf
constexpr uint32_t Phase_mesh = 2;  // at a node of an instance's mesh tree

template <bool Any>
__global__ void __launch_bounds__(Block2) trace_instances_refill(Tlas_view t, uint32_t* spill,
                                                                 uint32_t total, uint32_t* counters,
                                                                 mi_ray const* rays, uint32_t num_rays,
                                                                 uint32_t const* num_rays_dev,
                                                                 void* out) {
    __shared__ uint32_t lds[2 * Lds_entries * Block2];
    Pair_stack          stack;
    stack.lds   = lds;
    stack.spill = spill;
    stack.gid   = blockIdx.x * blockDim.x + threadIdx.x;
    stack.total = total;
    stack.sp    = 0;

    if (num_rays_dev) {
        num_rays = min(num_rays, *num_rays_dev);
    }
    uint32_t const lane    = threadIdx.x & 63;
    uint64_t const lt_mask = (uint64_t(1) << lane) - 1;

    Claims claims;
    claims.init();
    bool     active    = false;
    bool     exhausted = false;
    uint32_t pool_next = 0;
    uint32_t pool_end  = 0;

    uint32_t      r     = 0;
    uint32_t      phase = Phase_top;
    Iray          w, o;
    uint32_t      data = 0, meta = 0;
    uint32_t      leaf_i = 0, leaf_end = 0;
    uint32_t      inst   = Null;
    float4 const* mnodes = nullptr;
    float4 const* mtris  = nullptr;
    uint32_t      instance = Null, index = Null, slot = Null;
    float         u = 0.f, v = 0.f;

    auto finish = [&](bool any) {
        if (Any) {
            static_cast<uint32_t*>(out)[r] = any ? 1u : 0u;
        } else {
            mi_instance_hit h;
            h.t        = w.max_t;
            h.u        = Null != index ? u : 0.f;
            h.v        = Null != index ? v : 0.f;
            h.triangle = index;
            h.instance = instance;
            h.slot     = slot;
            h.pad[0] = h.pad[1] = 0;
            static_cast<mi_instance_hit*>(out)[r] = h;
        }
        active = false;
    };

    for (;;) {
        uint64_t       idle   = __ballot(!active);
        uint32_t const n_idle = uint32_t(__popcll(idle));
        if ((n_idle >= Refill_lanes || 64 == n_idle) && !(exhausted && pool_next >= pool_end)) {
            while (idle) {
                if (pool_next >= pool_end) {
                    if (exhausted) {
                        break;
                    }
                    uint32_t const leader = uint32_t(__ffsll(static_cast<unsigned long long>(idle))) - 1;
                    if (!claims.next(counters, num_rays, Claim_batch, lane, leader, pool_next, pool_end)) {
                        exhausted = true;
                        break;
                    }
                }
                uint32_t const avail = pool_end - pool_next;
                uint32_t const want  = uint32_t(__popcll(idle));
                uint32_t const take  = avail < want ? avail : want;
                bool const     mine  = (idle >> lane) & 1u;
                uint32_t const rank  = uint32_t(__popcll(idle & lt_mask));
                if (mine && rank < take) {
                    r                = pool_next + rank;
                    float4 const* rp = reinterpret_cast<float4 const*>(rays) + 2 * size_t(r);
                    float4 const  ro = rp[0];
                    float4 const  rd = rp[1];
                    w.o              = {ro.x, ro.y, ro.z};
                    w.min_t          = ro.w;
                    w.d              = {rd.x, rd.y, rd.z};
                    w.max_t          = rd.w;
                    set_inverse(w);
                    instance = index = slot = Null;
                    u = v    = 0.f;
                    stack.sp = 0;
                    phase    = Phase_top;
                    active   = false;
                    if (t.num_nodes) {
                        float4 const rlo = t.nodes[0];
                        float4 const rhi = t.nodes[1];
                        data             = __float_as_uint(rlo.w);
                        meta             = __float_as_uint(rhi.w);
                        active           = dev::slab(rlo, rhi, w.o, w.inv, w.min_t, w.max_t);
                    }
                    if (!active) {
                        finish(false);
                    }
                }
                pool_next += take;
                idle = __ballot(mine && rank >= take);
            }
        }

        if (0 == __ballot(active)) {
            if (exhausted && pool_next >= pool_end) {
                break;
            }
            continue;
        }
        if (!active) {
            continue;
        }

        if (Phase_top == phase) {
            if (0 == node_num(meta)) {
                if (!pair_step(t.nodes, stack, w, data, meta)) {
                    finish(false);
                }
            } else {
                leaf_i   = data;
                leaf_end = data + node_num(meta);
                phase    = Phase_leaf;
            }
        } else if (Phase_leaf == phase) {
            if (leaf_i < leaf_end) {
                uint32_t const         p  = __float_as_uint(t.slots[3 * size_t(leaf_i)].w);
                Instance_record const& ir = t.instances[p];
                ++leaf_i;
                float4 const blo = make_float4(ir.lo[0], ir.lo[1], ir.lo[2], 0.f);
                float4 const bhi = make_float4(ir.hi[0], ir.hi[1], ir.hi[2], 0.f);
                // Prop::intersect: the prop's world box first
                if (dev::slab(blo, bhi, w.o, w.inv, w.min_t, w.max_t)) {
                    o.o     = xform_point(ir.world_to_object, w.o);
                    o.d     = xform_vector(ir.world_to_object, w.d);
                    o.min_t = w.min_t;
                    o.max_t = w.max_t;
                    set_inverse(o);
                    Mesh_view const mv  = t.meshes[ir.mesh];
                    float4 const    rlo = mv.nodes[0];
                    float4 const    rhi = mv.nodes[1];
                    if (dev::slab(rlo, rhi, o.o, o.inv, o.min_t, o.max_t)) {
                        stack.push(Sentinel_word, 0.f);
                        mnodes = mv.nodes;
                        mtris  = mv.tris;
                        inst   = p;
                        data   = __float_as_uint(rlo.w);
                        meta   = __float_as_uint(rhi.w);
                        phase  = Phase_mesh;
                    }
                }
            } else if (pop_pending(stack, w, data, meta)) {
                phase = Phase_top;
            } else {
                finish(false);
            }
        } else {
            uint32_t const num  = node_num(meta);
            bool           live = true;
            if (0 == num) {
                live = pair_step(mnodes, stack, o, data, meta);
            } else {
                bool occluded = false;
                for (uint32_t i = data, end = data + num; i < end; ++i) {
                    float4 const ta = mtris[3 * size_t(i) + 0];
                    float4 const t1 = mtris[3 * size_t(i) + 1];
                    float4 const t2 = mtris[3 * size_t(i) + 2];
                    float        th, tu, tv;
                    if (dev::triangle({ta.x, ta.y, ta.z}, {t1.x, t1.y, t1.z}, {t2.x, t2.y, t2.z}, o.o,
                                      o.d, o.min_t, o.max_t, th, tu, tv)) {
                        if (Any) {
                            occluded = true;
                            break;
                        }
                        o.max_t  = th;
                        u        = tu;
                        v        = tv;
                        index    = __float_as_uint(ta.w);
                        slot     = i;
                        instance = inst;
                    }
                }
                if (Any && occluded) {
                    finish(true);
                    continue;
                }
                live = pop_pending(stack, o, data, meta);
            }
            if (!live) {
                // the mesh's sentinel is popped: back to the leaf's next instance
                w.max_t = o.max_t;
                phase   = Phase_leaf;
            }
        }
    }
}

__global__ void reset_claims(uint32_t* counters) {
    if (threadIdx.x < Claim_parts) {
        counters[threadIdx.x * Claim_stride] = 0;
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
    (void)hipFree(spill);
    (void)hipFree(claims);
    spill       = nullptr;
    claims      = nullptr;
    spill_total = 0;
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

    // the two-word stack traversal: every tree's node words must pack, and
    // the deepest stack beyond its LDS part spills to a per-thread area
    bool packs = out.tree.num_nodes < (1u << Data_bits) && out.tree.num_slots < (1u << Data_bits);
    for (uint32_t k = 0; k < num_meshes; ++k) {
        packs = packs && meshes[k]->num_nodes < (1u << Data_bits) &&
                meshes[k]->num_slots < (1u << Data_bits);
    }
    if (packs) {
        int device = 0, cus = 0;
        e = hipGetDevice(&device);
        if (!e) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
        uint32_t const total  = uint32_t(std::max(cus, 1)) * Blocks_per_cu * Block2;
        uint32_t const deeper = out.max_stack > Lds_entries ? out.max_stack - Lds_entries : 0;
        if (!e) e = hipMalloc(&out.spill, std::max<size_t>(1, size_t(2) * deeper * total) * sizeof(uint32_t));
        if (!e) e = hipMalloc(&out.claims, Counter_words * sizeof(uint32_t));
        if (e) {
            out.release();
            return fail(MI_ERR_HIP, "instances: stack area: %s", hipGetErrorString(e));
        }
        out.spill_total = total;
    }
    return MI_OK;
}

int32_t trace_instances_launch(Tlas const& t, uint32_t num_cus, mi_ray const* rays, uint32_t n,
                               uint32_t const* n_dev, void* out, bool any_hit, hipStream_t stream) {
    if (0 == n) {
        return MI_OK;
    }
    Tlas_view const v = tlas_view(t);
    if (t.spill) {
        uint64_t const wanted = (uint64_t(n) + Block2 - 1) / Block2;
        uint32_t const blocks = uint32_t(std::min<uint64_t>(wanted, t.spill_total / Block2));
        hipLaunchKernelGGL(reset_claims, dim3(1), dim3(64), 0, stream, t.claims);
        if (any_hit) {
            hipLaunchKernelGGL(trace_instances_refill<true>, dim3(blocks), dim3(Block2), 0, stream, v,
                               t.spill, t.spill_total, t.claims, rays, n, n_dev, out);
        } else {
            hipLaunchKernelGGL(trace_instances_refill<false>, dim3(blocks), dim3(Block2), 0, stream, v,
                               t.spill, t.spill_total, t.claims, rays, n, n_dev, out);
        }
        if (hipError_t const e = hipGetLastError(); e) {
            return fail(MI_ERR_HIP, "instance trace: %s", hipGetErrorString(e));
        }
        return MI_OK;
    }
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
