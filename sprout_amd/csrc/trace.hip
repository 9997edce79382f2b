// Closest-hit and any-hit traversal of sprout's binary SBVH on gfx950.
//
// Replaces scene::shape::triangle::bvh::Tree::intersect / intersect_p
// (core/scene/shape/triangle/bvh/triangle_bvh_tree.cpp:41-90,145-185).
//
// Tree layout in HBM (built by bvh_build.hip or uploaded):
//   nodes: sprout's 32-byte scene::bvh::Node, two float4 per node
//          {min.xyz, children_or_data}{max.xyz, axis | num_indices << 8}.
//          The array starts 32 bytes into a 64-byte aligned allocation, so
//          every sibling pair (children a, a + 1 with a odd) is one aligned
//          64-byte line segment.
//   tris:  48 bytes per tree slot {a.xyz, input triangle id}{b - a}{c - a}.
//          The edge vectors are the same rounded differences the reference
//          computes per test, so the intersection arithmetic is unchanged.
//
// Two loops, both bit-identical to the reference in result:
//   * trace_pairs: one thread per ray over a persistent grid.  At an inner
//     node both children are tested from one 64-byte fetch; the far child
//     (by the sign of inv_dir on the node's axis, as the reference) is pushed
//     only when its box is hit, and re-tested with the current max_t when
//     popped.  The box test only gets stricter as max_t shrinks, so the
//     triangles are tested in exactly the reference's order.
//   * trace_reference: the reference's loop verbatim (fetch, test, push far,
//     descend near), counting node fetches and triangle tests per ray; this
//     gives the algorithmic bytes per ray.
//   * trace_refill(_packed): persistent waves that refill idle lanes with new
//     rays; the packed variant's stack entries carry the far child's node
//     word and box entry distance, so a pop needs no node fetch.
// The traversal stack lives in LDS (Lds_entries per lane, laid out
// entry-major so a wave's 64 lanes hit 64 different banks) and spills to a
// per-thread global area beyond that, sized from the tree's depth.

#include "common.hpp"
#include "device_math.hpp"
#include "claims.hpp"
#include "trace.hpp"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>

namespace mi {

namespace {

using dev::F3;
using dev::Null;

constexpr uint32_t Max_lds_entries = 32;

struct Stack {
    uint32_t* lds;          // lds_entries * block_size words, entry-major
    uint32_t* spill;        // spill entries * total_threads words, entry-major
    uint32_t  lds_entries;  // <= Max_lds_entries
    uint32_t  lane;         // threadIdx.x
    uint32_t  block_size;
    uint32_t  gid;
    uint32_t  total_threads;
    uint32_t  sp;

    __device__ void init(uint32_t* lds_base, uint32_t* spill_base, uint32_t entries,
                         uint32_t threads) {
        lds           = lds_base;
        spill         = spill_base;
        lds_entries   = entries;
        lane          = threadIdx.x;
        block_size    = blockDim.x;
        gid           = blockIdx.x * blockDim.x + threadIdx.x;
        total_threads = threads;
        sp            = 0;
    }

    __device__ void push(uint32_t v) {
        if (sp < lds_entries) {
            lds[sp * block_size + lane] = v;
        } else {
            spill[size_t(sp - lds_entries) * total_threads + gid] = v;
        }
        ++sp;
    }

    __device__ uint32_t pop() {
        --sp;
        if (sp < lds_entries) {
            return lds[sp * block_size + lane];
        }
        return spill[size_t(sp - lds_entries) * total_threads + gid];
    }
};

struct Launch {
    float4 const* nodes;
    uint4 const*  qpairs;     // kernel 4: compressed sibling pairs
    float         qorigin[3];
    float         qscale[3];
    float4 const* tris;
    mi_ray const* rays;
    uint32_t      num_rays;
    uint32_t const* num_rays_dev;  // if set, trace_refill reads the ray count here
    void*         out;
    uint32_t*     counts;
    uint32_t*     spill;
    uint32_t      total_threads;
    uint32_t      lds_entries;
    uint32_t*     counter;     // claim counters of trace_refill: Claim_parts words, Claim_stride apart
    uint32_t      refill_lanes;
    uint32_t      claim_batch;
    uint32_t const* order;    // kernels 3 and 4: the k-th claimed ray is order[k] (or k)
};

struct Ray {
    F3    o;
    F3    d;
    F3    inv;
    float min_t;
    float max_t;
    uint32_t sign_bits;  // bit a set when inv_dir[a] < 0 (base/math/simd.inl:230-236)
};

__device__ __forceinline__ Ray load_ray(mi_ray const* rays, uint32_t r) {
    float4 const* p  = reinterpret_cast<float4 const*>(rays) + 2 * size_t(r);
    float4 const  ro = p[0];
    float4 const  rd = p[1];

    Ray ray;
    ray.o     = {ro.x, ro.y, ro.z};
    ray.min_t = ro.w;
    ray.d     = {rd.x, rd.y, rd.z};
    ray.max_t = rd.w;
    ray.inv   = {dev::reciprocal_lane(rd.x), dev::reciprocal_lane(rd.y), dev::reciprocal_lane(rd.z)};
    ray.sign_bits = (ray.inv.x < 0.f ? 1u : 0u) | (ray.inv.y < 0.f ? 2u : 0u) |
                    (ray.inv.z < 0.f ? 4u : 0u);
    return ray;
}

__device__ __forceinline__ uint32_t node_num(uint32_t meta) {
    return (meta >> 8) & 0xFFu;
}

__device__ __forceinline__ uint32_t node_axis(uint32_t meta) {
    return meta & 0xFFu;
}

// Tests the triangles of a leaf in slot order.  Closest: every accepted hit
// shrinks max_t (ties accepted, so the later equal-t triangle wins).
// Any: returns true on the first hit.
template <bool Any>
__device__ __forceinline__ bool leaf(float4 const* __restrict__ tris, uint32_t start, uint32_t num,
                                     Ray& ray, uint32_t& index, float& u, float& v,
                                     uint32_t& tests) {
    for (uint32_t i = start, end = start + num; i < end; ++i) {
        ++tests;

        float4 const ta = tris[3 * size_t(i) + 0];
        float4 const t1 = tris[3 * size_t(i) + 1];
        float4 const t2 = tris[3 * size_t(i) + 2];

        float t, tu, tv;
        if (dev::triangle({ta.x, ta.y, ta.z}, {t1.x, t1.y, t1.z}, {t2.x, t2.y, t2.z}, ray.o, ray.d,
                          ray.min_t, ray.max_t, t, tu, tv)) {
            if (Any) {
                return true;
            }
            ray.max_t = t;
            u         = tu;
            v         = tv;
            index     = __float_as_uint(ta.w);
        }
    }
    return false;
}

template <bool Any>
__global__ void trace_pairs(Launch p) {
    extern __shared__ uint32_t lds_stack[];

    float4 const* __restrict__ nodes = p.nodes;
    float4 const* __restrict__ tris  = p.tris;
    mi_ray const* __restrict__ rays  = p.rays;
    uint32_t const num_rays          = p.num_rays;
    uint32_t const total_threads     = p.total_threads;
    void* const    out               = p.out;

    Stack stack;
    stack.init(lds_stack, p.spill, p.lds_entries, total_threads);

    for (uint32_t r = stack.gid; r < num_rays; r += total_threads) {
        Ray ray = load_ray(rays, r);

        uint32_t index = Null;
        float    u     = 0.f;
        float    v     = 0.f;
        bool     hit_p = false;

        stack.sp = 0;

        float4 const rlo = nodes[0];
        float4 const rhi = nodes[1];

        uint32_t data = __float_as_uint(rlo.w);
        uint32_t meta = __float_as_uint(rhi.w);

        bool live = dev::slab(rlo, rhi, ray.o, ray.inv, ray.min_t, ray.max_t);

        while (live) {
            uint32_t const num = node_num(meta);

            if (0 == num) {
                uint32_t const a = data;

                float4 const alo = nodes[2 * size_t(a) + 0];
                float4 const ahi = nodes[2 * size_t(a) + 1];
                float4 const blo = nodes[2 * size_t(a) + 2];
                float4 const bhi = nodes[2 * size_t(a) + 3];

                bool const ha = dev::slab(alo, ahi, ray.o, ray.inv, ray.min_t, ray.max_t);
                bool const hb = dev::slab(blo, bhi, ray.o, ray.inv, ray.min_t, ray.max_t);

                bool const flip = (ray.sign_bits >> node_axis(meta)) & 1u;

                bool const hn = flip ? hb : ha;
                bool const hf = flip ? ha : hb;

                if (hn | hf) {
                    if (hn & hf) {
                        stack.push(flip ? a : a + 1);
                    }
                    bool const take_b = hn ? flip : !flip;
                    data = __float_as_uint(take_b ? blo.w : alo.w);
                    meta = __float_as_uint(take_b ? bhi.w : ahi.w);
                    continue;
                }
            } else {
                uint32_t tests = 0;
                if (leaf<Any>(tris, data, num, ray, index, u, v, tests)) {
                    hit_p = true;
                    break;
                }
            }

            // pop until a pending far child still passes the box test
            live = false;
            while (stack.sp > 0) {
                uint32_t const m   = stack.pop();
                float4 const   mlo = nodes[2 * size_t(m) + 0];
                float4 const   mhi = nodes[2 * size_t(m) + 1];
                if (dev::slab(mlo, mhi, ray.o, ray.inv, ray.min_t, ray.max_t)) {
                    data = __float_as_uint(mlo.w);
                    meta = __float_as_uint(mhi.w);
                    live = true;
                    break;
                }
            }
        }

        if (Any) {
            static_cast<uint32_t*>(out)[r] = hit_p ? 1u : 0u;
        } else {
            float4 h;
            h.x = ray.max_t;
            h.y = Null != index ? u : 0.f;
            h.z = Null != index ? v : 0.f;
            h.w = __uint_as_float(index);
            reinterpret_cast<float4*>(out)[r] = h;
        }
    }
}

template <bool Any>
__global__ void trace_reference(Launch p) {
    extern __shared__ uint32_t lds_stack[];

    float4 const* __restrict__ nodes = p.nodes;
    float4 const* __restrict__ tris  = p.tris;
    mi_ray const* __restrict__ rays  = p.rays;
    uint32_t const num_rays          = p.num_rays;
    uint32_t const total_threads     = p.total_threads;
    void* const    out               = p.out;
    uint32_t* const counts           = p.counts;

    Stack stack;
    stack.init(lds_stack, p.spill, p.lds_entries, total_threads);

    for (uint32_t r = stack.gid; r < num_rays; r += total_threads) {
        Ray ray = load_ray(rays, r);

        uint32_t index = Null;
        float    u     = 0.f;
        float    v     = 0.f;
        bool     hit_p = false;

        uint32_t num_nodes = 0;
        uint32_t num_tris  = 0;

        stack.sp = 0;
        stack.push(Null);

        uint32_t n = 0;
        while (Null != n) {
            float4 const lo = nodes[2 * size_t(n) + 0];
            float4 const hi = nodes[2 * size_t(n) + 1];
            ++num_nodes;

            if (dev::slab(lo, hi, ray.o, ray.inv, ray.min_t, ray.max_t)) {
                uint32_t const data = __float_as_uint(lo.w);
                uint32_t const meta = __float_as_uint(hi.w);
                uint32_t const num  = node_num(meta);

                if (0 == num) {
                    if (0 == ((ray.sign_bits >> node_axis(meta)) & 1u)) {
                        stack.push(data + 1);
                        n = data;
                    } else {
                        stack.push(data);
                        n = data + 1;
                    }
                    continue;
                }

                // intersect_p stops at the first hit: only the tests made count
                if (leaf<Any>(tris, data, num, ray, index, u, v, num_tris)) {
                    hit_p = true;
                    break;
                }
            }

            n = stack.pop();
        }

        if (Any) {
            static_cast<uint32_t*>(out)[r] = hit_p ? 1u : 0u;
        } else {
            float4 h;
            h.x = ray.max_t;
            h.y = Null != index ? u : 0.f;
            h.z = Null != index ? v : 0.f;
            h.w = __uint_as_float(index);
            reinterpret_cast<float4*>(out)[r] = h;
        }

        counts[2 * size_t(r) + 0] = num_nodes;
        counts[2 * size_t(r) + 1] = num_tris;
    }
}


__device__ __forceinline__ void write_result(Launch const& p, uint32_t r, bool any, bool hit_p,
                                             float max_t, uint32_t index, float u, float v) {
    if (any) {
        static_cast<uint32_t*>(p.out)[r] = hit_p ? 1u : 0u;
    } else {
        float4 h;
        h.x = max_t;
        h.y = Null != index ? u : 0.f;
        h.z = Null != index ? v : 0.f;
        h.w = __uint_as_float(index);
        reinterpret_cast<float4*>(p.out)[r] = h;
    }
}

// Pops pending far children until one still passes the box test with the
// current max_t; returns false when the stack runs empty (the ray is done).
__device__ __forceinline__ bool pop_next(float4 const* __restrict__ nodes, Stack& stack,
                                         Ray const& ray, uint32_t& data, uint32_t& meta) {
    while (stack.sp > 0) {
        uint32_t const m   = stack.pop();
        float4 const   mlo = nodes[2 * size_t(m) + 0];
        float4 const   mhi = nodes[2 * size_t(m) + 1];
        if (dev::slab(mlo, mhi, ray.o, ray.inv, ray.min_t, ray.max_t)) {
            data = __float_as_uint(mlo.w);
            meta = __float_as_uint(mhi.w);
            return true;
        }
    }
    return false;
}

// One inner node: both children from one 64-byte fetch, near/far by the
// sign of inv_dir on the node's axis (triangle_bvh_tree.cpp:64-70).
// Returns false when the ray is done.
__device__ __forceinline__ bool inner_step(float4 const* __restrict__ nodes, Stack& stack,
                                           Ray const& ray, uint32_t& data, uint32_t& meta) {
    uint32_t const a = data;

    float4 const alo = nodes[2 * size_t(a) + 0];
    float4 const ahi = nodes[2 * size_t(a) + 1];
    float4 const blo = nodes[2 * size_t(a) + 2];
    float4 const bhi = nodes[2 * size_t(a) + 3];

    bool const ha = dev::slab(alo, ahi, ray.o, ray.inv, ray.min_t, ray.max_t);
    bool const hb = dev::slab(blo, bhi, ray.o, ray.inv, ray.min_t, ray.max_t);

    bool const flip = (ray.sign_bits >> node_axis(meta)) & 1u;
    bool const hn   = flip ? hb : ha;
    bool const hf   = flip ? ha : hb;

    if (hn | hf) {
        if (hn & hf) {
            stack.push(flip ? a : a + 1);
        }
        bool const take_b = hn ? flip : !flip;
        data              = __float_as_uint(take_b ? blo.w : alo.w);
        meta              = __float_as_uint(take_b ? bhi.w : ahi.w);
        return true;
    }
    return pop_next(nodes, stack, ray, data, meta);
}

// Persistent traversal with per-lane ray refill: a lane whose ray is done
// takes the next unclaimed ray, so the 64 lanes of a wave stay busy while
// rays of very different lengths finish.  Each wave claims rays in batches
// of claim_batch from the global counter (one atomic per batch) and hands
// them out from a wave-uniform pool once at least refill_lanes lanes are
// idle (or all are).  WhileWhile separates inner-node steps from leaf steps
// (all lanes descend until every active lane sits at a leaf); the visiting
// order of each ray is the same either way.
template <bool Any, bool WhileWhile>
__global__ void trace_refill(Launch p) {
    extern __shared__ uint32_t lds_stack[];

    float4 const* __restrict__ nodes = p.nodes;
    float4 const* __restrict__ tris  = p.tris;

    Stack stack;
    stack.init(lds_stack, p.spill, p.lds_entries, p.total_threads);

    uint32_t const lane     = threadIdx.x & 63;
    uint64_t const lt_mask  = (uint64_t(1) << lane) - 1;
    uint32_t const num_rays = p.num_rays_dev ? *p.num_rays_dev : p.num_rays;

    Claims   claims;
    claims.init();
    bool     active    = false;
    bool     exhausted = false;  // the global counter is past the last ray (wave-uniform)
    uint32_t pool_next = 0;      // wave-uniform pool of claimed rays [pool_next, pool_end)
    uint32_t pool_end  = 0;
    uint32_t r         = 0;
    Ray      ray;
    uint32_t data  = 0;
    uint32_t meta  = 0;
    uint32_t index = Null;
    float    u     = 0.f;
    float    v     = 0.f;

    for (;;) {
        uint64_t idle = __ballot(!active);
        uint32_t const n_idle = uint32_t(__popcll(idle));

        if ((n_idle >= p.refill_lanes || 64 == n_idle) && !(exhausted && pool_next >= pool_end)) {
            while (idle) {
                if (pool_next >= pool_end) {
                    if (exhausted) {
                        break;
                    }
                    uint32_t const leader =
                        uint32_t(__ffsll(static_cast<unsigned long long>(idle))) - 1;
                    if (!claims.next(p.counter, num_rays, p.claim_batch, lane, leader, pool_next,
                                     pool_end)) {
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
                    r        = pool_next + rank;
                    ray      = load_ray(p.rays, r);
                    index    = Null;
                    u        = 0.f;
                    v        = 0.f;
                    stack.sp = 0;

                    float4 const rlo = nodes[0];
                    float4 const rhi = nodes[1];
                    if (dev::slab(rlo, rhi, ray.o, ray.inv, ray.min_t, ray.max_t)) {
                        data   = __float_as_uint(rlo.w);
                        meta   = __float_as_uint(rhi.w);
                        active = true;
                    } else {
                        write_result(p, r, Any, false, ray.max_t, Null, 0.f, 0.f);
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

        if (WhileWhile) {
            while (__ballot(active && 0 == node_num(meta))) {
                if (active && 0 == node_num(meta)) {
                    if (!inner_step(nodes, stack, ray, data, meta)) {
                        write_result(p, r, Any, false, ray.max_t, index, u, v);
                        active = false;
                    }
                }
            }
        }

        if (active) {
            uint32_t const num = node_num(meta);
            bool           live;
            bool           hit_p = false;
            if (0 == num) {
                live = inner_step(nodes, stack, ray, data, meta);
            } else {
                uint32_t tests = 0;
                hit_p          = leaf<Any>(tris, data, num, ray, index, u, v, tests);
                live           = !hit_p && pop_next(nodes, stack, ray, data, meta);
            }
            if (!live) {
                write_result(p, r, Any, hit_p, ray.max_t, index, u, v);
                active = false;
            }
        }
    }
}

// Two-word stack entries {far child's node word, its box entry distance}:
// the node word packs the child's data (< 2^22) with its axis and primitive
// count, so a popped child continues without fetching its node again, and a
// popped child whose entry distance exceeds the current max_t is skipped: the
// reference's box test of that child at this max_t is exactly max_t >= tmin
// (slab_entry), the rest of it held when the child was pushed.
constexpr uint32_t Packed_data_bits = 22;

__device__ __forceinline__ uint32_t pack_node(uint32_t data, uint32_t meta) {
    // the axis of an inner node is 0..2; a leaf's is never read
    return data | ((node_axis(meta) & 3u) << Packed_data_bits) | (node_num(meta) << 24);
}

__device__ __forceinline__ void unpack_node(uint32_t e, uint32_t& data, uint32_t& meta) {
    data = e & ((1u << Packed_data_bits) - 1u);
    meta = ((e >> Packed_data_bits) & 3u) | ((e >> 24) << 8);
}

struct Stack2 {
    uint32_t* lds;          // 2 * lds_entries * block_size words, entry-major
    uint32_t* spill;        // 2 * spill entries * total_threads words, entry-major
    uint32_t  lds_entries;
    uint32_t  lane;
    uint32_t  block_size;
    uint32_t  gid;
    uint32_t  total_threads;
    uint32_t  sp;

    __device__ void init(uint32_t* lds_base, uint32_t* spill_base, uint32_t entries,
                         uint32_t threads) {
        lds           = lds_base;
        spill         = spill_base;
        lds_entries   = entries;
        lane          = threadIdx.x;
        block_size    = blockDim.x;
        gid           = blockIdx.x * blockDim.x + threadIdx.x;
        total_threads = threads;
        sp            = 0;
    }

    __device__ void push(uint32_t node, float tmin) {
        if (sp < lds_entries) {
            lds[(2 * sp) * block_size + lane]     = node;
            lds[(2 * sp + 1) * block_size + lane] = __float_as_uint(tmin);
        } else {
            size_t const k                                 = 2 * size_t(sp - lds_entries);
            spill[k * total_threads + gid]       = node;
            spill[(k + 1) * total_threads + gid] = __float_as_uint(tmin);
        }
        ++sp;
    }

    __device__ void pop(uint32_t& node, float& tmin) {
        --sp;
        if (sp < lds_entries) {
            node = lds[(2 * sp) * block_size + lane];
            tmin = __uint_as_float(lds[(2 * sp + 1) * block_size + lane]);
        } else {
            size_t const k = 2 * size_t(sp - lds_entries);
            node           = spill[k * total_threads + gid];
            tmin           = __uint_as_float(spill[(k + 1) * total_threads + gid]);
        }
    }
};

__device__ __forceinline__ bool pop_next2(Stack2& stack, Ray const& ray, uint32_t& data,
                                          uint32_t& meta) {
    while (stack.sp > 0) {
        uint32_t e;
        float    tmin;
        stack.pop(e, tmin);
        if (ray.max_t >= tmin) {
            unpack_node(e, data, meta);
            return true;
        }
    }
    return false;
}

__device__ __forceinline__ bool inner_step2(float4 const* __restrict__ nodes, Stack2& stack,
                                            Ray const& ray, uint32_t& data, uint32_t& meta) {
    uint32_t const a = data;

    float4 const alo = nodes[2 * size_t(a) + 0];
    float4 const ahi = nodes[2 * size_t(a) + 1];
    float4 const blo = nodes[2 * size_t(a) + 2];
    float4 const bhi = nodes[2 * size_t(a) + 3];

    float      ta, tb;
    bool const ha = dev::slab_entry(alo, ahi, ray.o, ray.inv, ray.min_t, ray.max_t, ta);
    bool const hb = dev::slab_entry(blo, bhi, ray.o, ray.inv, ray.min_t, ray.max_t, tb);

    bool const flip = (ray.sign_bits >> node_axis(meta)) & 1u;
    bool const hn   = flip ? hb : ha;
    bool const hf   = flip ? ha : hb;

    if (hn | hf) {
        if (hn & hf) {
            uint32_t const fd = __float_as_uint(flip ? alo.w : blo.w);
            uint32_t const fm = __float_as_uint(flip ? ahi.w : bhi.w);
            stack.push(pack_node(fd, fm), flip ? ta : tb);
        }
        bool const take_b = hn ? flip : !flip;
        data              = __float_as_uint(take_b ? blo.w : alo.w);
        meta              = __float_as_uint(take_b ? bhi.w : ahi.w);
        return true;
    }
    return pop_next2(stack, ray, data, meta);
}

// A sibling pair from its 32-byte compressed record: each box coordinate is
// origin + q * scale on a 16-bit grid over the root box, the lower corners
// rounded down and the upper ones up when the record was made, so each box
// holds the original; the words are pack_node's.
__device__ __forceinline__ float dequant(uint32_t q, float scale, float origin) {
    return __builtin_fmaf(float(q), scale, origin);
}

__device__ __forceinline__ bool inner_step_q(Launch const& p, Stack2& stack, Ray const& ray,
                                             uint32_t& data, uint32_t& meta) {
    uint32_t const a   = data;
    uint4 const*   rec = p.qpairs + 2 * size_t((a - 1) >> 1);
    uint4 const    r0  = rec[0];  // lo_a.xyz hi_a.xyz lo_b.xy as 16-bit pairs
    uint4 const    r1  = rec[1];  // lo_b.z hi_b.xyz, word_a, word_b
    float const    ox = p.qorigin[0], oy = p.qorigin[1], oz = p.qorigin[2];
    float const    sx = p.qscale[0], sy = p.qscale[1], sz = p.qscale[2];
    float4 const alo = make_float4(dequant(r0.x & 0xFFFFu, sx, ox), dequant(r0.x >> 16, sy, oy),
                                   dequant(r0.y & 0xFFFFu, sz, oz), 0.f);
    float4 const ahi = make_float4(dequant(r0.y >> 16, sx, ox), dequant(r0.z & 0xFFFFu, sy, oy),
                                   dequant(r0.z >> 16, sz, oz), 0.f);
    float4 const blo = make_float4(dequant(r0.w & 0xFFFFu, sx, ox), dequant(r0.w >> 16, sy, oy),
                                   dequant(r1.x & 0xFFFFu, sz, oz), 0.f);
    float4 const bhi = make_float4(dequant(r1.x >> 16, sx, ox), dequant(r1.y & 0xFFFFu, sy, oy),
                                   dequant(r1.y >> 16, sz, oz), 0.f);

    float      ta, tb;
    bool const ha = dev::slab_entry(alo, ahi, ray.o, ray.inv, ray.min_t, ray.max_t, ta);
    bool const hb = dev::slab_entry(blo, bhi, ray.o, ray.inv, ray.min_t, ray.max_t, tb);

    bool const flip = (ray.sign_bits >> node_axis(meta)) & 1u;
    bool const hn   = flip ? hb : ha;
    bool const hf   = flip ? ha : hb;

    if (hn | hf) {
        if (hn & hf) {
            stack.push(flip ? r1.z : r1.w, flip ? ta : tb);
        }
        bool const take_b = hn ? flip : !flip;
        unpack_node(take_b ? r1.w : r1.z, data, meta);
        return true;
    }
    return pop_next2(stack, ray, data, meta);
}

// trace_refill<Any, true> with the two-word stack (trees of < 2^22 nodes and
// slots); the same rays in the same order, the same results.  Quant: inner
// nodes from the compressed pairs (kernel 4); a box of a pair holds the
// original, so every node the exact boxes admit is visited, in the same
// near / far order (the axis and the ray's sign decide it), and the
// triangles are tested in the same relative order.
template <bool Any, bool Quant>
__global__ void trace_refill_packed(Launch p) {
    extern __shared__ uint32_t lds_stack[];

    float4 const* __restrict__ nodes = p.nodes;
    float4 const* __restrict__ tris  = p.tris;

    Stack2 stack;
    stack.init(lds_stack, p.spill, p.lds_entries, p.total_threads);

    uint32_t const lane     = threadIdx.x & 63;
    uint64_t const lt_mask  = (uint64_t(1) << lane) - 1;
    uint32_t const num_rays = p.num_rays_dev ? *p.num_rays_dev : p.num_rays;

    Claims   claims;
    claims.init();
    bool     active    = false;
    bool     exhausted = false;
    uint32_t pool_next = 0;
    uint32_t pool_end  = 0;
    uint32_t r         = 0;
    Ray      ray;
    uint32_t data  = 0;
    uint32_t meta  = 0;
    uint32_t index = Null;
    float    u     = 0.f;
    float    v     = 0.f;

    for (;;) {
        uint64_t idle = __ballot(!active);
        uint32_t const n_idle = uint32_t(__popcll(idle));

        if ((n_idle >= p.refill_lanes || 64 == n_idle) && !(exhausted && pool_next >= pool_end)) {
            while (idle) {
                if (pool_next >= pool_end) {
                    if (exhausted) {
                        break;
                    }
                    uint32_t const leader =
                        uint32_t(__ffsll(static_cast<unsigned long long>(idle))) - 1;
                    if (!claims.next(p.counter, num_rays, p.claim_batch, lane, leader, pool_next,
                                     pool_end)) {
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
                    r        = p.order ? p.order[pool_next + rank] : pool_next + rank;
                    ray      = load_ray(p.rays, r);
                    index    = Null;
                    u        = 0.f;
                    v        = 0.f;
                    stack.sp = 0;

                    float4 const rlo = nodes[0];
                    float4 const rhi = nodes[1];
                    if (dev::slab(rlo, rhi, ray.o, ray.inv, ray.min_t, ray.max_t)) {
                        data   = __float_as_uint(rlo.w);
                        meta   = __float_as_uint(rhi.w);
                        active = true;
                    } else {
                        write_result(p, r, Any, false, ray.max_t, Null, 0.f, 0.f);
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

        while (__ballot(active && 0 == node_num(meta))) {
            if (active && 0 == node_num(meta)) {
                bool const more = Quant ? inner_step_q(p, stack, ray, data, meta)
                                        : inner_step2(nodes, stack, ray, data, meta);
                if (!more) {
                    write_result(p, r, Any, false, ray.max_t, index, u, v);
                    active = false;
                }
            }
        }

        if (active) {
            uint32_t const num = node_num(meta);
            bool           live;
            bool           hit_p = false;
            if (0 == num) {
                live = inner_step2(nodes, stack, ray, data, meta);
            } else {
                uint32_t tests = 0;
                hit_p          = leaf<Any>(tris, data, num, ray, index, u, v, tests);
                live           = !hit_p && pop_next2(stack, ray, data, meta);
            }
            if (!live) {
                write_result(p, r, Any, hit_p, ray.max_t, index, u, v);
                active = false;
            }
        }
    }
}

// One compressed sibling pair per thread (children 2j + 1, 2j + 2): 16-bit
// grid coordinates, lower corners the largest grid value at or below the
// coordinate, upper ones the smallest at or above (checked with the
// traversal's own dequantization).
__device__ uint32_t quant_lo(float x, float scale, float origin) {
    float    g = floorf((x - origin) / scale);
    uint32_t q = g <= 0.f ? 0u : (g >= 65535.f ? 65535u : uint32_t(g));
    while (q > 0 && dequant(q, scale, origin) > x) {
        --q;
    }
    return q;
}

__device__ uint32_t quant_hi(float x, float scale, float origin) {
    float    g = ceilf((x - origin) / scale);
    uint32_t q = g <= 0.f ? 0u : (g >= 65535.f ? 65535u : uint32_t(g));
    while (q < 65535u && dequant(q, scale, origin) < x) {
        ++q;
    }
    return q;
}

// A pair record serves the inner node whose children start at 2j + 1; a
// tree whose children start at an even index cannot use them (*odd_only
// then gets 0).
__global__ void quantize_pairs(float4 const* nodes, uint32_t num_pairs, float ox, float oy,
                               float oz, float sx, float sy, float sz, uint4* out,
                               uint32_t* odd_only) {
    uint32_t const j = blockIdx.x * blockDim.x + threadIdx.x;
    if (0 == j) {
        float4 const rhi = nodes[1];
        if (0 == node_num(__float_as_uint(rhi.w)) && 0 == (__float_as_uint(nodes[0].w) & 1u)) {
            atomicAnd(odd_only, 0u);
        }
    }
    if (j >= num_pairs) {
        return;
    }
    size_t const a   = 2 * size_t(j) + 1;
    float4 const alo = nodes[2 * a + 0];
    float4 const ahi = nodes[2 * a + 1];
    float4 const blo = nodes[2 * a + 2];
    float4 const bhi = nodes[2 * a + 3];
    if ((0 == node_num(__float_as_uint(ahi.w)) && 0 == (__float_as_uint(alo.w) & 1u)) ||
        (0 == node_num(__float_as_uint(bhi.w)) && 0 == (__float_as_uint(blo.w) & 1u))) {
        atomicAnd(odd_only, 0u);
    }
    uint32_t const q[12] = {quant_lo(alo.x, sx, ox), quant_lo(alo.y, sy, oy), quant_lo(alo.z, sz, oz),
                            quant_hi(ahi.x, sx, ox), quant_hi(ahi.y, sy, oy), quant_hi(ahi.z, sz, oz),
                            quant_lo(blo.x, sx, ox), quant_lo(blo.y, sy, oy), quant_lo(blo.z, sz, oz),
                            quant_hi(bhi.x, sx, ox), quant_hi(bhi.y, sy, oy), quant_hi(bhi.z, sz, oz)};
    uint4 r0, r1;
    r0.x = q[0] | (q[1] << 16);
    r0.y = q[2] | (q[3] << 16);
    r0.z = q[4] | (q[5] << 16);
    r0.w = q[6] | (q[7] << 16);
    r1.x = q[8] | (q[9] << 16);
    r1.y = q[10] | (q[11] << 16);
    r1.z = pack_node(__float_as_uint(alo.w), __float_as_uint(ahi.w));
    r1.w = pack_node(__float_as_uint(blo.w), __float_as_uint(bhi.w));
    out[2 * size_t(j) + 0] = r0;
    out[2 * size_t(j) + 1] = r1;
}

// Sort key of each ray: its direction octant over a 27-bit Morton code of
// its origin on a 512^3 grid over the tree's box (origins outside clamp).
__device__ __forceinline__ uint32_t spread3(uint32_t x) {
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__global__ void ray_keys(mi_ray const* rays, uint32_t n, float lx, float ly, float lz, float sx,
                         float sy, float sz, uint32_t* keys, uint32_t* order) {
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    float4 const* p  = reinterpret_cast<float4 const*>(rays) + 2 * size_t(i);
    float4 const  ro = p[0];
    float4 const  rd = p[1];
    auto cell = [](float x, float lo, float s) {
        float const c = (x - lo) * s;
        return c >= 511.f ? 511u : (c > 0.f ? uint32_t(c) : 0u);
    };
    uint32_t const oct = (rd.x < 0.f ? 4u : 0u) | (rd.y < 0.f ? 2u : 0u) | (rd.z < 0.f ? 1u : 0u);
    uint32_t const m   = (spread3(cell(ro.x, lx, sx)) << 2) | (spread3(cell(ro.y, ly, sy)) << 1) |
                       spread3(cell(ro.z, lz, sz));
    keys[i]  = (oct << 27) | m;
    order[i] = i;
}

__global__ void reset_counter(uint32_t* counter) {
    if (threadIdx.x < Claim_parts) {
        counter[threadIdx.x * Claim_stride] = 0;
    }
}

}  // namespace

namespace {

// The two-word stack holds node words of trees below 2^22 nodes and slots.
bool packable(DeviceBvh const& bvh) {
    return bvh.num_nodes < (1u << Packed_data_bits) && bvh.num_slots < (1u << Packed_data_bits);
}

// The compressed pairs of kernel 4 (made once per tree): a grid of 65 535
// steps per axis over the root box, its scale grown until the last step
// reaches the box's upper corner.
int32_t ensure_quantized(DeviceBvh const& bvh, hipStream_t stream) {
    if (bvh.qpairs || bvh.qfailed || bvh.num_nodes < 3) {
        return MI_OK;
    }
    uint32_t const num_pairs = (bvh.num_nodes - 1) / 2;
    for (int k = 0; k < 3; ++k) {
        float const lo = bvh.aabb_min[k];
        float const hi = bvh.aabb_max[k];
        float       sc = (hi - lo) / 65535.f;
        if (!(sc > 0.f)) {
            sc = 1.17549435e-38f;
        }
        while (std::fma(65535.f, sc, lo) < hi) {
            sc = std::nextafter(sc, 3.40282347e+38f);
        }
        bvh.qorigin[k] = lo;
        bvh.qscale[k]  = sc;
    }
    uint4*     q    = nullptr;
    uint32_t*  flag = nullptr;
    hipError_t e    = hipMalloc(&q, size_t(num_pairs) * 2 * sizeof(uint4));
    if (!e) e = hipMalloc(&flag, sizeof(uint32_t));
    uint32_t odd = 1;
    if (!e) e = hipMemcpyAsync(flag, &odd, sizeof(uint32_t), hipMemcpyHostToDevice, stream);
    if (e) {
        (void)hipFree(q);
        (void)hipFree(flag);
        return fail(MI_ERR_HIP, "compressed tree of %u pairs: %s", num_pairs, hipGetErrorString(e));
    }
    hipLaunchKernelGGL(quantize_pairs, dim3((num_pairs + 255) / 256), dim3(256), 0, stream,
                       bvh.nodes, num_pairs, bvh.qorigin[0], bvh.qorigin[1], bvh.qorigin[2],
                       bvh.qscale[0], bvh.qscale[1], bvh.qscale[2], q, flag);
    e = hipMemcpyAsync(&odd, flag, sizeof(uint32_t), hipMemcpyDeviceToHost, stream);
    if (!e) e = hipStreamSynchronize(stream);
    (void)hipFree(flag);
    if (e) {
        (void)hipFree(q);
        return fail(MI_ERR_HIP, "compressed tree: %s", hipGetErrorString(e));
    }
    if (!odd) {
        (void)hipFree(q);
        bvh.qfailed = true;  // kernel 3 serves this tree
        return MI_OK;
    }
    bvh.qpairs = q;
    return MI_OK;
}

// Stack entries per lane in LDS (at most what a workgroup's 160 KB hold) and
// in the global spill area, and words per entry.
struct Stack_layout {
    uint32_t words;
    uint32_t lds_entries;
    uint32_t spill_entries;
};

Stack_layout stack_layout(DeviceBvh const& bvh, TraceResources const& res, bool counts) {
    uint32_t const entries = stack_entries(bvh.max_depth);
    Stack_layout   l;
    l.words                = (!counts && 3 <= res.kernel && packable(bvh)) ? 2 : 1;
    uint32_t const lds_cap = res.lds_entries ? res.lds_entries : Max_lds_entries;
    uint32_t const fit     = (160u * 1024u) / (l.words * res.block_size * 4u);
    uint32_t const cap     = lds_cap < fit ? lds_cap : fit;
    l.lds_entries          = entries < cap ? entries : cap;
    l.spill_entries        = entries - l.lds_entries;
    return l;
}

// A fixed scratch area: the ray counter and a stack spill area of a given
// capacity, which the launch fits its persistent grid into.
struct Fixed_scratch {
    uint32_t* counter;
    uint32_t* spill;
    size_t    spill_words;
};

// num_rays_dev set: the refill kernel reads the ray count on the device and
// num_rays is only its upper bound (it sizes the grid and the stack spill).
// fixed set: the launch allocates nothing and uses fixed's counter and spill
// area, with as many workgroups (at most the usual grid) as the area holds.
int32_t launch(DeviceBvh const& bvh, TraceResources& res, mi_ray const* rays, uint32_t num_rays,
               uint32_t const* num_rays_dev, void* out, uint32_t* counts, bool any_hit,
               hipStream_t stream, Fixed_scratch const* fixed = nullptr) {
    if (0 == num_rays) {
        return MI_OK;
    }

    uint32_t const entries = stack_entries(bvh.max_depth);
    if (entries > Reference_stack_size) {
        return fail(MI_ERR_STACK,
                    "tree depth %u needs %u stack entries, more than the reference's %u",
                    bvh.max_depth, entries, Reference_stack_size);
    }

    uint32_t const block  = res.block_size;
    uint32_t       blocks = res.grid_blocks(num_rays);

    Stack_layout const sl          = stack_layout(bvh, res, counts != nullptr);
    uint32_t const     lds_entries = sl.lds_entries;
    bool const         packed      = 2 == sl.words;
    size_t const       lane_words  = size_t(sl.words) * sl.spill_entries;
    uint32_t*          spill       = nullptr;
    uint32_t*          counter     = nullptr;
    if (fixed) {
        if (lane_words) {
            size_t const fit = fixed->spill_words / (lane_words * block);
            if (0 == fit) {
                return fail(MI_ERR_ARGS,
                            "tree depth %u spills %zu bytes of traversal stack per workgroup of %u, "
                            "more than the fixed scratch's %zu bytes",
                            bvh.max_depth, lane_words * block * sizeof(uint32_t), block,
                            fixed->spill_words * sizeof(uint32_t));
            }
            blocks = uint32_t(std::min<size_t>(blocks, fit));
        }
        spill   = fixed->spill;
        counter = fixed->counter;
    } else {
        if (int32_t const st = res.reserve_spill(lane_words * blocks * block); st) {
            return st;
        }
        spill   = res.spill;
        counter = res.counter;
    }
    uint32_t const threads = blocks * block;

    size_t const lds_bytes = size_t(sl.words) * lds_entries * block * sizeof(uint32_t);

    // (a launch that must not allocate uses the compressed pairs only once made)
    if (packed && 4 == res.kernel && bvh.num_nodes >= 3 && !fixed) {
        if (int32_t const st = ensure_quantized(bvh, stream); st) {
            return st;
        }
    }
    bool const quant = packed && 4 == res.kernel && nullptr != bvh.qpairs;

    // the rays' tracing order (kernels 3 and 4, host count, allocating launches)
    uint32_t const* order = nullptr;
    if (res.sort_rays && packed && !counts && !num_rays_dev && !fixed && num_rays > 1) {
        if (num_rays > res.sort_capacity) {
            (void)hipFree(res.sort_keys);
            (void)hipFree(res.sort_temp);
            res.sort_keys       = nullptr;
            res.sort_temp       = nullptr;
            res.sort_capacity   = 0;
            res.sort_temp_bytes = 0;
            size_t     bytes = 0;
            hipError_t e     = hipcub::DeviceRadixSort::SortPairs(
                nullptr, bytes, static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr), int(num_rays), 0,
                30, stream);
            if (!e) e = hipMalloc(&res.sort_keys, 4 * size_t(num_rays) * sizeof(uint32_t));
            if (!e) e = hipMalloc(&res.sort_temp, std::max<size_t>(bytes, 1));
            if (e) {
                return fail(MI_ERR_HIP, "ray order of %u rays: %s", num_rays, hipGetErrorString(e));
            }
            res.sort_capacity   = num_rays;
            res.sort_temp_bytes = std::max<size_t>(bytes, 1);
        }
        uint32_t* const keys_in   = res.sort_keys;
        uint32_t* const keys_out  = keys_in + res.sort_capacity;
        uint32_t* const order_in  = keys_out + res.sort_capacity;
        uint32_t* const order_out = order_in + res.sort_capacity;
        float           scale[3];
        for (int k = 0; k < 3; ++k) {
            float const ext = bvh.aabb_max[k] - bvh.aabb_min[k];
            scale[k]        = ext > 0.f ? 512.f / ext : 0.f;
        }
        hipLaunchKernelGGL(ray_keys, dim3((num_rays + 255) / 256), dim3(256), 0, stream, rays,
                           num_rays, bvh.aabb_min[0], bvh.aabb_min[1], bvh.aabb_min[2], scale[0],
                           scale[1], scale[2], keys_in, order_in);
        size_t     bytes = res.sort_temp_bytes;
        hipError_t e     = hipcub::DeviceRadixSort::SortPairs(res.sort_temp, bytes, keys_in,
                                                              keys_out, order_in, order_out,
                                                              int(num_rays), 0, 30, stream);
        if (e) {
            return fail(MI_ERR_HIP, "ray order: %s", hipGetErrorString(e));
        }
        order = order_out;
    }

    Launch p;
    p.order         = order;
    p.nodes         = bvh.nodes;
    p.qpairs        = bvh.qpairs;
    for (int k = 0; k < 3; ++k) {
        p.qorigin[k] = bvh.qorigin[k];
        p.qscale[k]  = bvh.qscale[k];
    }
    p.tris          = bvh.tris;
    p.rays          = rays;
    p.num_rays      = num_rays;
    p.num_rays_dev  = num_rays_dev;
    p.out           = out;
    p.counts        = counts;
    p.spill         = spill;
    p.total_threads = threads;
    p.lds_entries   = lds_entries;
    p.counter       = counter;
    p.refill_lanes  = res.refill_lanes;
    p.claim_batch   = res.claim_batch;

    if (counts) {
        if (any_hit) {
            hipLaunchKernelGGL(trace_reference<true>, dim3(blocks), dim3(block), lds_bytes, stream, p);
        } else {
            hipLaunchKernelGGL(trace_reference<false>, dim3(blocks), dim3(block), lds_bytes, stream, p);
        }
    } else if (packed) {
        hipLaunchKernelGGL(reset_counter, dim3(1), dim3(64), 0, stream, counter);
        using Kernel = void (*)(Launch);
        Kernel const kern = any_hit ? (quant ? trace_refill_packed<true, true> : trace_refill_packed<true, false>)
                                    : (quant ? trace_refill_packed<false, true> : trace_refill_packed<false, false>);
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(block), lds_bytes, stream, p);
    } else if (res.kernel != 1 || num_rays_dev) {
        hipLaunchKernelGGL(reset_counter, dim3(1), dim3(64), 0, stream, counter);
        using Kernel    = void (*)(Launch);
        bool const ww   = 2 <= res.kernel;
        Kernel     kern = any_hit ? (ww ? trace_refill<true, true> : trace_refill<true, false>)
                                  : (ww ? trace_refill<false, true> : trace_refill<false, false>);
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(block), lds_bytes, stream, p);
    } else if (any_hit) {
        hipLaunchKernelGGL(trace_pairs<true>, dim3(blocks), dim3(block), lds_bytes, stream, p);
    } else {
        hipLaunchKernelGGL(trace_pairs<false>, dim3(blocks), dim3(block), lds_bytes, stream, p);
    }

    if (hipError_t const e = hipGetLastError(); e != hipSuccess) {
        return fail(MI_ERR_HIP, "trace launch: %s", hipGetErrorString(e));
    }
    return MI_OK;
}

}  // namespace

int32_t prepare_trace(DeviceBvh const& bvh, TraceResources& res, uint32_t max_rays) {
    uint32_t const entries = stack_entries(bvh.max_depth);
    if (entries > Reference_stack_size) {
        return fail(MI_ERR_STACK,
                    "tree depth %u needs %u stack entries, more than the reference's %u",
                    bvh.max_depth, entries, Reference_stack_size);
    }
    uint32_t const     threads = res.grid_blocks(max_rays) * res.block_size;
    Stack_layout const sl      = stack_layout(bvh, res, false);
    if (4 == res.kernel && 2 == sl.words) {
        // made before any capture: a launch inside a graph allocates nothing
        if (int32_t const st = ensure_quantized(bvh, nullptr); st) {
            return st;
        }
    }
    return res.reserve_spill(size_t(sl.words) * sl.spill_entries * threads);
}

int32_t launch_trace(DeviceBvh const& bvh, TraceResources& res, mi_ray const* rays,
                     uint32_t num_rays, void* out, uint32_t* counts, bool any_hit,
                     hipStream_t stream) {
    return launch(bvh, res, rays, num_rays, nullptr, out, counts, any_hit, stream);
}

int32_t launch_trace_device_count(DeviceBvh const& bvh, TraceResources& res, mi_ray const* rays,
                                  uint32_t const* num_rays, uint32_t max_rays, void* out,
                                  bool any_hit, hipStream_t stream) {
    return launch(bvh, res, rays, max_rays, num_rays, out, nullptr, any_hit, stream);
}

// ---- closest hit without allocation ----

namespace {

// Scratch of the traversal that allocates nothing: the ray claim counters
// (Claim_parts lines of 256 bytes), then the stack spill area.  Uninitialised device data:
// every GPU the library runs on holds its own copy until the process ends.
// In host code `scratch` names a host-side stand-in; kernels get its device
// address, looked up per device at set-up.
constexpr size_t Noalloc_scratch_bytes = size_t(32) << 20;
constexpr size_t Noalloc_counter_words = size_t(Claim_parts) * Claim_stride;
static __device__ __align__(256) uint32_t scratch[Noalloc_scratch_bytes / sizeof(uint32_t)];

constexpr int Noalloc_max_devices = 64;
uint32_t*     scratch_address[Noalloc_max_devices] = {};

}  // namespace

int32_t noalloc_scratch_setup(int device) {
    if (device < 0 || device >= Noalloc_max_devices) {
        return fail(MI_ERR_ARGS, "fixed traversal scratch: device %d out of range", device);
    }
    if (scratch_address[device]) {
        return MI_OK;
    }
    void* ptr = nullptr;
    if (hipError_t const e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)); e != hipSuccess) {
        return fail(MI_ERR_HIP, "fixed traversal scratch on device %d: %s", device,
                    hipGetErrorString(e));
    }
    scratch_address[device] = static_cast<uint32_t*>(ptr);
    return MI_OK;
}

int32_t launch_trace_noalloc(DeviceBvh const& bvh, TraceResources& res, int device,
                             mi_ray const* rays, uint32_t num_rays, mi_hit* hits,
                             hipStream_t stream) {
    uint32_t* const base =
        (device >= 0 && device < Noalloc_max_devices) ? scratch_address[device] : nullptr;
    if (!base) {
        return fail(MI_ERR_NO_CONTEXT, "fixed traversal scratch not set up on device %d", device);
    }
    Fixed_scratch const fixed{base, base + Noalloc_counter_words,
                              Noalloc_scratch_bytes / sizeof(uint32_t) - Noalloc_counter_words};
    return launch(bvh, res, rays, num_rays, nullptr, hits, nullptr, false, stream, &fixed);
}

}  // namespace mi
