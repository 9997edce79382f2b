// GPU construction of sprout's triangle SBVH.
//
// Produces the tree scene::shape::triangle::bvh::Builder_SAH produces
// (triangle_bvh_builder_sah.cpp:21-67 over scene_bvh_builder_base.cpp:26-290
// and scene_bvh_split_candidate.inl:49-151): the same split decisions, the
// same leaves in the same order, serialized depth-first with sibling pairs
// as serialize() does (triangle_bvh_builder_sah.cpp:129-164).
//
// The reference recurses depth-first on a thread pool; the tree it builds
// depends only on each node's reference list, so here all nodes of one
// depth are processed together, level by level:
//
//   plan      per node: leaf (<= max_primitives), sweep node (<= sweep
//             threshold: candidates at every reference's max bound) or grid
//             node (3 centre planes + num_slices planes per axis, object and,
//             above the spatial depth, spatial); counts and offsets.
//   evaluate  grid nodes: one workgroup per 1024-reference chunk, each wave
//             folds a subset of candidates over the chunk and merges its
//             counts and boxes into per-candidate accumulators with integer
//             atomics (min/max of order-preserving float encodings, so the
//             result does not depend on order).  Sweep nodes: one workgroup
//             per node, references in LDS.
//   select    per node: SAH cost of every candidate with the reference's
//             rounding, first minimum, exhausted test, leaf/split decision.
//   move      per chunk: stable partition of the references into the two
//             children (straddling references clipped to the plane), leaf
//             references appended to the leaf list in order.
//
// Finally sub-tree sizes (bottom-up) give every node its depth-first
// position (top-down) and the nodes and slots are scattered into place.
//
// Departures, all in inputs the reference cannot handle either: a node of
// more than 255 references that cannot be split fails with
// MI_ERR_UNSPLITTABLE (the reference logs and leaves the node
// uninitialized); a grid node whose box has zero extent gets no slice
// candidates (the reference converts NaN to an integer step count).  Bounds
// equal up to the sign of zero may differ in that sign.

#include "bvh_build.hpp"
#include "device_math.hpp"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

namespace mi {

namespace {

constexpr uint32_t Chunk      = 1024;  // references per workgroup in evaluate / move
constexpr uint32_t Kind_leaf  = 1;
constexpr uint32_t Kind_sweep = 2;
constexpr uint32_t Kind_grid  = 3;
constexpr uint32_t Kind_split = 4;
constexpr uint32_t Kind_error = 5;

constexpr float Flt_max = 3.40282347e+38f;

// scene::bvh::Reference: {min.xyz, primitive}{max.xyz, pad}
struct Ref {
    float    lo[3];
    uint32_t prim;
    float    hi[3];
    uint32_t pad;
};
static_assert(sizeof(Ref) == 32);

struct Task {
    uint32_t begin;
    uint32_t count;
    uint32_t node;
    uint32_t depth;
    float    lo[3];
    float    hi[3];
};

struct Cand {
    float    d;
    uint32_t axis;
    uint32_t spatial;
};

// per-candidate accumulator: counts and boxes as order-preserving keys
struct Acc {
    uint32_t n0;
    uint32_t n1;
    uint32_t used;
    uint32_t pad;
    uint32_t lo0[3];
    uint32_t hi0[3];
    uint32_t lo1[3];
    uint32_t hi1[3];
};

struct Decision {
    uint32_t kind;
    uint32_t axis;
    uint32_t spatial;
    uint32_t n0;
    uint32_t n1;
    float    d;
    float    lo0[3];
    float    hi0[3];
    float    lo1[3];
    float    hi1[3];
};

struct Chunk_rec {
    uint32_t task;
    uint32_t begin;
    uint32_t end;
};

// sprout's Node layout (scene_bvh_node.hpp:44-59)
struct Node {
    float    lo[3];
    uint32_t data;
    float    hi[3];
    uint32_t meta;  // axis | num_indices << 8
};
static_assert(sizeof(Node) == 32);

struct Settings {
    uint32_t num_slices;
    uint32_t sweep_threshold;
    uint32_t max_primitives;
    uint32_t spatial_split_threshold;
};

// ---- float helpers with the reference's semantics --------------------------

// SSE minps/maxps: min(a, b) = a < b ? a : b
__device__ __forceinline__ float smin(float a, float b) {
    return a < b ? a : b;
}

__device__ __forceinline__ float smax(float a, float b) {
    return a > b ? a : b;
}

// std::min / std::max
__device__ __forceinline__ float stdmin(float a, float b) {
    return b < a ? b : a;
}

__device__ __forceinline__ float stdmax(float a, float b) {
    return a < b ? b : a;
}

__device__ __forceinline__ uint32_t key(float f) {
    uint32_t const b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ float unkey(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// base/math/aabb.inl:39-42 as the reference build contracts it
__device__ __forceinline__ float surface_area(float const lo[3], float const hi[3]) {
    float const d0 = hi[0] - lo[0];
    float const d1 = hi[1] - lo[1];
    float const d2 = hi[2] - lo[2];
    float const s  = __builtin_fmaf(d1, d2, __builtin_fmaf(d0, d1, d0 * d2));
    return s + s;
}

// Split_candidate::evaluate cost (scene_bvh_split_candidate.inl:109-118)
__device__ __forceinline__ float sah_cost(uint32_t n0, uint32_t n1, uint32_t n, float const lo0[3],
                                          float const hi0[3], float const lo1[3],
                                          float const hi1[3], float aabb_surface_area) {
    if (0 == n0 || 0 == n1) {
        return 2.f + float(n);
    }
    float const sa0      = surface_area(lo0, hi0);
    float const sa1      = surface_area(lo1, hi1);
    float const weight_1 = float(n1) * sa1;
    float const weights  = __builtin_fmaf(float(n0), sa0, weight_1);
    float const k        = float(n0 + n1 - n);
    return __builtin_fmaf(k, 0.125f, weights / aabb_surface_area + 2.f);
}

__device__ __forceinline__ uint32_t index_max_component(float const v[3]) {
    if (v[0] > v[1]) {
        return v[0] > v[2] ? 0 : 2;
    }
    return v[1] > v[2] ? 1 : 2;
}

// Number of grid slices on axis a (scene_bvh_builder_base.cpp:119-129).
__device__ __forceinline__ uint32_t grid_steps(float extent_a, float step, uint32_t num_slices) {
    float const q = extent_a / step;
    // a zero-extent box gives NaN; a count beyond num_slices + 1 cannot come
    // from a finite box (extent_a <= extent of the longest axis)
    if (!(q >= 0.f) || q > float(num_slices + 1)) {
        return 0;
    }
    return uint32_t(ceilf(q));
}

__device__ __forceinline__ bool covers(float const alo[3], float const ahi[3], float const blo[3],
                                       float const bhi[3]) {
    return alo[0] <= blo[0] && alo[1] <= blo[1] && alo[2] <= blo[2] && ahi[0] >= bhi[0] &&
           ahi[1] >= bhi[1] && ahi[2] >= bhi[2];
}

// ---- kernels ----------------------------------------------------------------

__global__ void make_refs(float const* __restrict__ positions, uint32_t stride,
                          uint32_t num_vertices, uint32_t const* __restrict__ indices,
                          uint32_t num_triangles, Ref* __restrict__ refs,
                          uint32_t* __restrict__ root_keys, uint32_t* __restrict__ error) {
    uint32_t const t = blockIdx.x * blockDim.x + threadIdx.x;

    float lo[3] = {Flt_max, Flt_max, Flt_max};
    float hi[3] = {-Flt_max, -Flt_max, -Flt_max};

    if (t < num_triangles) {
        uint32_t ia = indices[3 * size_t(t) + 0];
        uint32_t ib = indices[3 * size_t(t) + 1];
        uint32_t ic = indices[3 * size_t(t) + 2];
        if (ia >= num_vertices || ib >= num_vertices || ic >= num_vertices) {
            atomicOr(error, 2u);
            ia = ib = ic = 0;
        }
        float const* pa = positions + size_t(ia) * stride;
        float const* pb = positions + size_t(ib) * stride;
        float const* pc = positions + size_t(ic) * stride;

        Ref r;
        for (int i = 0; i < 3; ++i) {
            // triangle_bvh_helper.cpp:15-21: min(a, min(b, c))
            r.lo[i] = smin(pa[i], smin(pb[i], pc[i]));
            r.hi[i] = smax(pa[i], smax(pb[i], pc[i]));
            lo[i]   = r.lo[i];
            hi[i]   = r.hi[i];
        }
        r.prim  = t;
        r.pad   = 0;
        refs[t] = r;
    }

    // root box: wave reduction, then one atomic per wave and coordinate
    for (int i = 0; i < 3; ++i) {
        for (int off = 32; off > 0; off >>= 1) {
            lo[i] = fminf(lo[i], __shfl_xor(lo[i], off));
            hi[i] = fmaxf(hi[i], __shfl_xor(hi[i], off));
        }
    }
    if (0 == (threadIdx.x & 63)) {
        for (int i = 0; i < 3; ++i) {
            atomicMin(&root_keys[i], key(lo[i]));
            atomicMax(&root_keys[3 + i], key(hi[i]));
        }
    }
}

__global__ void init_root(uint32_t const* __restrict__ root_keys, uint32_t num_triangles,
                          Task* __restrict__ tasks) {
    Task t;
    t.begin = 0;
    t.count = num_triangles;
    t.node  = 0;
    t.depth = 0;
    for (int i = 0; i < 3; ++i) {
        t.lo[i] = unkey(root_keys[i]);
        t.hi[i] = unkey(root_keys[3 + i]);
    }
    tasks[0] = t;
}

// Per task: kind, grid candidate count, chunk counts.
__global__ void plan(Task const* __restrict__ tasks, uint32_t num_tasks, Settings s,
                     uint32_t* __restrict__ kind, uint32_t* __restrict__ cand_count,
                     uint32_t* __restrict__ grid_chunks, uint32_t* __restrict__ all_chunks) {
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num_tasks) {
        return;
    }
    Task const     t = tasks[i];
    uint32_t const n = t.count;
    uint32_t const c = (n + Chunk - 1) / Chunk;

    all_chunks[i] = c;

    if (n <= s.max_primitives) {
        kind[i]        = Kind_leaf;
        cand_count[i]  = 0;
        grid_chunks[i] = 0;
        return;
    }
    if (n <= s.sweep_threshold) {
        kind[i]        = Kind_sweep;
        cand_count[i]  = 0;
        grid_chunks[i] = 0;
        return;
    }

    float extent[3];
    for (int a = 0; a < 3; ++a) {
        extent[a] = t.hi[a] - t.lo[a];
    }
    uint32_t const la   = index_max_component(extent);
    float const    step = extent[la] / float(s.num_slices);

    uint32_t const per_slice = t.depth < s.spatial_split_threshold ? 2 : 1;

    uint32_t count = 3;
    for (int a = 0; a < 3; ++a) {
        uint32_t const steps = grid_steps(extent[a], step, s.num_slices);
        if (steps > 1) {
            count += (steps - 1) * per_slice;
        }
    }

    kind[i]        = Kind_grid;
    cand_count[i]  = count;
    grid_chunks[i] = c;
}

// Writes the grid candidates of each grid task in the reference's order
// (scene_bvh_builder_base.cpp:101-140) and clears their accumulators.
__global__ void make_candidates(Task const* __restrict__ tasks, uint32_t num_tasks, Settings s,
                                uint32_t const* __restrict__ kind,
                                uint32_t const* __restrict__ cand_offset, Cand* __restrict__ cands,
                                Acc* __restrict__ accs) {
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num_tasks || Kind_grid != kind[i]) {
        return;
    }
    Task const t = tasks[i];

    float position[3];
    float extent[3];
    for (int a = 0; a < 3; ++a) {
        position[a] = (t.lo[a] + t.hi[a]) * 0.5f;
        extent[a]   = t.hi[a] - t.lo[a];
    }

    uint32_t o = cand_offset[i];

    for (uint32_t a = 0; a < 3; ++a) {
        cands[o++] = {position[a], a, 1};
    }

    uint32_t const la      = index_max_component(extent);
    float const    step    = extent[la] / float(s.num_slices);
    bool const     spatial = t.depth < s.spatial_split_threshold;

    for (uint32_t a = 0; a < 3; ++a) {
        uint32_t const steps  = grid_steps(extent[a], step, s.num_slices);
        float const    step_a = extent[a] / float(steps);
        for (uint32_t k = 1; k < steps; ++k) {
            float const d = __builtin_fmaf(float(k), step_a, t.lo[a]);
            cands[o++]    = {d, a, 0};
            if (spatial) {
                cands[o++] = {d, a, 1};
            }
        }
    }

    Acc empty;
    empty.n0   = 0;
    empty.n1   = 0;
    empty.used = 0;
    empty.pad  = 0;
    for (int a = 0; a < 3; ++a) {
        empty.lo0[a] = key(Flt_max);
        empty.hi0[a] = key(-Flt_max);
        empty.lo1[a] = key(Flt_max);
        empty.hi1[a] = key(-Flt_max);
    }
    for (uint32_t c = cand_offset[i]; c < o; ++c) {
        accs[c] = empty;
    }
}

__global__ void make_chunks(Task const* __restrict__ tasks, uint32_t num_tasks,
                            uint32_t const* __restrict__ kind,
                            uint32_t const* __restrict__ grid_chunk_offset,
                            uint32_t const* __restrict__ all_chunk_offset,
                            Chunk_rec* __restrict__ grid_list, Chunk_rec* __restrict__ all_list) {
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num_tasks) {
        return;
    }
    Task const     t = tasks[i];
    uint32_t const c = (t.count + Chunk - 1) / Chunk;
    for (uint32_t k = 0; k < c; ++k) {
        uint32_t const  b = t.begin + k * Chunk;
        uint32_t const  e = min(b + Chunk, t.begin + t.count);
        Chunk_rec const r = {i, b, e};
        all_list[all_chunk_offset[i] + k] = r;
        if (Kind_grid == kind[i]) {
            grid_list[grid_chunk_offset[i] + k] = r;
        }
    }
}

struct Partial {
    uint32_t n0, n1, used;
    float    lo0[3], hi0[3], lo1[3], hi1[3];
};

__device__ __forceinline__ void partial_init(Partial& p) {
    p.n0 = p.n1 = p.used = 0;
    for (int a = 0; a < 3; ++a) {
        p.lo0[a] = Flt_max;
        p.hi0[a] = -Flt_max;
        p.lo1[a] = Flt_max;
        p.hi1[a] = -Flt_max;
    }
}

__device__ __forceinline__ void merge0(Partial& p, Ref const& r) {
    for (int a = 0; a < 3; ++a) {
        p.lo0[a] = smin(p.lo0[a], r.lo[a]);
        p.hi0[a] = smax(p.hi0[a], r.hi[a]);
    }
}

__device__ __forceinline__ void merge1(Partial& p, Ref const& r) {
    for (int a = 0; a < 3; ++a) {
        p.lo1[a] = smin(p.lo1[a], r.lo[a]);
        p.hi1[a] = smax(p.hi1[a], r.hi[a]);
    }
}

// Classification of one reference against a candidate
// (scene_bvh_split_candidate.inl:57-107).
__device__ __forceinline__ void classify(Partial& p, Ref const& r, uint32_t axis, float d,
                                         bool spatial) {
    if (r.hi[axis] < d) {
        ++p.n0;
        merge0(p, r);
    } else if (!spatial || !(r.lo[axis] < d)) {
        ++p.n1;
        merge1(p, r);
    } else {
        ++p.n0;
        ++p.n1;
        merge0(p, r);
        merge1(p, r);
        p.used = 1;
    }
}

__global__ void __launch_bounds__(256)
evaluate_grid(Chunk_rec const* __restrict__ chunks, Ref const* __restrict__ refs,
              uint32_t const* __restrict__ cand_offset, uint32_t const* __restrict__ cand_count,
              Cand const* __restrict__ cands, Acc* __restrict__ accs) {
    __shared__ Ref lds_refs[Chunk];

    Chunk_rec const ch = chunks[blockIdx.x];
    uint32_t const  n  = ch.end - ch.begin;

    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        lds_refs[i] = refs[ch.begin + i];
    }
    __syncthreads();

    uint32_t const lane    = threadIdx.x & 63;
    uint32_t const wave    = threadIdx.x >> 6;
    uint32_t const waves   = blockDim.x >> 6;
    uint32_t const c_begin = cand_offset[ch.task];
    uint32_t const c_count = cand_count[ch.task];

    for (uint32_t c = wave; c < c_count; c += waves) {
        Cand const cd = cands[c_begin + c];

        Partial p;
        partial_init(p);
        for (uint32_t i = lane; i < n; i += 64) {
            classify(p, lds_refs[i], cd.axis, cd.d, cd.spatial != 0);
        }

        for (int off = 32; off > 0; off >>= 1) {
            p.n0 += __shfl_xor(p.n0, off);
            p.n1 += __shfl_xor(p.n1, off);
            p.used |= __shfl_xor(p.used, off);
            for (int a = 0; a < 3; ++a) {
                p.lo0[a] = fminf(p.lo0[a], __shfl_xor(p.lo0[a], off));
                p.hi0[a] = fmaxf(p.hi0[a], __shfl_xor(p.hi0[a], off));
                p.lo1[a] = fminf(p.lo1[a], __shfl_xor(p.lo1[a], off));
                p.hi1[a] = fmaxf(p.hi1[a], __shfl_xor(p.hi1[a], off));
            }
        }

        if (0 == lane) {
            Acc* acc = &accs[c_begin + c];
            atomicAdd(&acc->n0, p.n0);
            atomicAdd(&acc->n1, p.n1);
            if (p.used) {
                atomicOr(&acc->used, 1u);
            }
            for (int a = 0; a < 3; ++a) {
                if (p.n0) {
                    atomicMin(&acc->lo0[a], key(p.lo0[a]));
                    atomicMax(&acc->hi0[a], key(p.hi0[a]));
                }
                if (p.n1) {
                    atomicMin(&acc->lo1[a], key(p.lo1[a]));
                    atomicMax(&acc->hi1[a], key(p.hi1[a]));
                }
            }
        }
    }
}

// Fills a decision from the winning candidate's counts and boxes
// (scene_bvh_builder_base.cpp:40-88 and the clipping of
// scene_bvh_split_candidate.inl:95-100).
__device__ void decide(Task const& t, uint32_t axis, float d, bool spatial_cand, Partial const& p,
                       float cost, Decision& out) {
    uint32_t const n = t.count;

    float lo0[3], hi0[3], lo1[3], hi1[3];
    for (int a = 0; a < 3; ++a) {
        lo0[a] = p.lo0[a];
        hi0[a] = p.hi0[a];
        lo1[a] = p.lo1[a];
        hi1[a] = p.hi1[a];
    }
    bool const spatial = spatial_cand && p.used;
    if (spatial) {
        hi0[axis] = stdmin(d, hi0[axis]);
        lo1[axis] = stdmax(d, lo1[axis]);
    }

    bool const exhausted = (covers(lo0, hi0, t.lo, t.hi) && n == p.n0) ||
                           (covers(lo1, hi1, t.lo, t.hi) && n == p.n1);

    uint32_t kind;
    if (n <= 0xFF && (float(n) <= cost || exhausted)) {
        kind = Kind_leaf;
    } else if (exhausted) {
        kind = Kind_error;
    } else if (n <= 0xFF && (0 == p.n0 || 0 == p.n1)) {
        kind = Kind_leaf;
    } else {
        kind = Kind_split;
    }

    out.kind    = kind;
    out.axis    = axis;
    out.spatial = spatial ? 1 : 0;
    out.n0      = p.n0;
    out.n1      = p.n1;
    out.d       = d;
    for (int a = 0; a < 3; ++a) {
        out.lo0[a] = lo0[a];
        out.hi0[a] = hi0[a];
        out.lo1[a] = lo1[a];
        out.hi1[a] = hi1[a];
    }
}

// First minimum of (cost, index) with the reference's scan semantics: a NaN
// cost never wins, except at index 0 where nothing can replace it.
struct Best {
    float    cost;
    uint32_t index;
};

__device__ __forceinline__ bool better(Best a, Best b) {
    // a strictly before b in the reference's sequential scan order
    if (0 == a.index || 0 == b.index) {
        bool const a_nan0 = (0 == a.index) && (a.cost != a.cost);
        bool const b_nan0 = (0 == b.index) && (b.cost != b.cost);
        if (a_nan0) {
            return true;
        }
        if (b_nan0) {
            return false;
        }
    }
    if (a.cost != a.cost) {
        return false;
    }
    if (b.cost != b.cost) {
        return true;
    }
    return a.cost < b.cost || (a.cost == b.cost && a.index < b.index);
}

__device__ Best block_best(Best mine, Best* lds) {
    for (int off = 32; off > 0; off >>= 1) {
        Best o;
        o.cost  = __shfl_xor(mine.cost, off);
        o.index = __shfl_xor(mine.index, off);
        if (better(o, mine)) {
            mine = o;
        }
    }
    uint32_t const wave  = threadIdx.x >> 6;
    uint32_t const waves = blockDim.x >> 6;
    if (0 == (threadIdx.x & 63)) {
        lds[wave] = mine;
    }
    __syncthreads();
    Best b = lds[0];
    for (uint32_t w = 1; w < waves; ++w) {
        if (better(lds[w], b)) {
            b = lds[w];
        }
    }
    __syncthreads();
    return b;
}

__global__ void __launch_bounds__(256)
select_grid(Task const* __restrict__ tasks, uint32_t const* __restrict__ kind,
            uint32_t const* __restrict__ cand_offset, uint32_t const* __restrict__ cand_count,
            Cand const* __restrict__ cands, Acc const* __restrict__ accs,
            Decision* __restrict__ decisions) {
    __shared__ Best lds[16];

    uint32_t const ti = blockIdx.x;
    if (Kind_grid != kind[ti]) {
        return;
    }
    Task const     t       = tasks[ti];
    uint32_t const c_begin = cand_offset[ti];
    uint32_t const c_count = cand_count[ti];
    float const    sa      = surface_area(t.lo, t.hi);

    Best mine = {__builtin_nanf(""), 0xFFFFFFFFu};
    for (uint32_t c = threadIdx.x; c < c_count; c += blockDim.x) {
        Acc const a = accs[c_begin + c];
        float lo0[3], hi0[3], lo1[3], hi1[3];
        for (int k = 0; k < 3; ++k) {
            lo0[k] = unkey(a.lo0[k]);
            hi0[k] = unkey(a.hi0[k]);
            lo1[k] = unkey(a.lo1[k]);
            hi1[k] = unkey(a.hi1[k]);
        }
        Cand const cd = cands[c_begin + c];
        if (cd.spatial && a.used) {
            hi0[cd.axis] = stdmin(cd.d, hi0[cd.axis]);
            lo1[cd.axis] = stdmax(cd.d, lo1[cd.axis]);
        }
        Best const b = {sah_cost(a.n0, a.n1, t.count, lo0, hi0, lo1, hi1, sa), c};
        if (better(b, mine)) {
            mine = b;
        }
    }

    Best const best = block_best(mine, lds);

    if (0 == threadIdx.x) {
        Acc const  a  = accs[c_begin + best.index];
        Cand const cd = cands[c_begin + best.index];
        Partial    p;
        p.n0   = a.n0;
        p.n1   = a.n1;
        p.used = a.used;
        for (int k = 0; k < 3; ++k) {
            p.lo0[k] = unkey(a.lo0[k]);
            p.hi0[k] = unkey(a.hi0[k]);
            p.lo1[k] = unkey(a.lo1[k]);
            p.hi1[k] = unkey(a.hi1[k]);
        }
        decide(t, cd.axis, cd.d, cd.spatial != 0, p, best.cost, decisions[ti]);
    }
}

// Sweep nodes (<= sweep threshold references): candidates are the three
// centre planes and every reference's max bound on each axis, evaluated
// over the node's references in LDS (scene_bvh_builder_base.cpp:106-114).
__global__ void __launch_bounds__(256)
select_sweep(Task const* __restrict__ tasks, uint32_t const* __restrict__ kind,
             Ref const* __restrict__ refs, Decision* __restrict__ decisions) {
    extern __shared__ Ref lds_refs[];
    __shared__ Best       lds[16];

    uint32_t const ti = blockIdx.x;
    if (Kind_sweep != kind[ti]) {
        return;
    }
    Task const     t = tasks[ti];
    uint32_t const n = t.count;

    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        lds_refs[i] = refs[t.begin + i];
    }
    __syncthreads();

    float const sa = surface_area(t.lo, t.hi);

    auto candidate = [&](uint32_t c, uint32_t& axis, float& d, bool& spatial) {
        if (c < 3) {
            axis    = c;
            d       = (t.lo[c] + t.hi[c]) * 0.5f;
            spatial = true;
        } else {
            uint32_t const r = (c - 3) / 3;
            axis             = (c - 3) % 3;
            d                = lds_refs[r].hi[axis];
            spatial          = false;
        }
    };

    uint32_t const c_count = 3 + 3 * n;

    Best mine = {__builtin_nanf(""), 0xFFFFFFFFu};
    for (uint32_t c = threadIdx.x; c < c_count; c += blockDim.x) {
        uint32_t axis;
        float    d;
        bool     spatial;
        candidate(c, axis, d, spatial);

        Partial p;
        partial_init(p);
        for (uint32_t i = 0; i < n; ++i) {
            classify(p, lds_refs[i], axis, d, spatial);
        }
        float lo0[3], hi0[3], lo1[3], hi1[3];
        for (int k = 0; k < 3; ++k) {
            lo0[k] = p.lo0[k];
            hi0[k] = p.hi0[k];
            lo1[k] = p.lo1[k];
            hi1[k] = p.hi1[k];
        }
        if (spatial && p.used) {
            hi0[axis] = stdmin(d, hi0[axis]);
            lo1[axis] = stdmax(d, lo1[axis]);
        }
        Best const b = {sah_cost(p.n0, p.n1, n, lo0, hi0, lo1, hi1, sa), c};
        if (better(b, mine)) {
            mine = b;
        }
    }

    Best const best = block_best(mine, lds);

    if (0 == threadIdx.x) {
        uint32_t axis;
        float    d;
        bool     spatial;
        candidate(best.index, axis, d, spatial);
        Partial p;
        partial_init(p);
        for (uint32_t i = 0; i < n; ++i) {
            classify(p, lds_refs[i], axis, d, spatial);
        }
        decide(t, axis, d, spatial, p, best.cost, decisions[ti]);
    }
}

// Per task: leaf size, split flag and children's reference count.
__global__ void outcome_counts(Task const* __restrict__ tasks, uint32_t num_tasks,
                               uint32_t const* __restrict__ kind,
                               Decision* __restrict__ decisions, uint32_t* __restrict__ leaf_count,
                               uint32_t* __restrict__ split_flag, uint32_t* __restrict__ child_refs,
                               uint32_t* __restrict__ error) {
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num_tasks) {
        return;
    }
    uint32_t const n = tasks[i].count;
    uint32_t       k = kind[i];
    if (Kind_leaf == k) {
        decisions[i].kind = Kind_leaf;
    } else {
        k = decisions[i].kind;
    }
    leaf_count[i] = Kind_leaf == k ? n : 0;
    split_flag[i] = Kind_split == k ? 1 : 0;
    child_refs[i] = Kind_split == k ? decisions[i].n0 + decisions[i].n1 : 0;
    if (Kind_error == k) {
        atomicOr(error, 1u);
    }
}

// Writes this level's build nodes and the next level's tasks.
__global__ void emit_nodes(Task const* __restrict__ tasks, uint32_t num_tasks,
                           Decision const* __restrict__ decisions,
                           uint32_t const* __restrict__ leaf_offset,
                           uint32_t const* __restrict__ split_rank,
                           uint32_t const* __restrict__ next_offset, uint32_t leaf_base,
                           uint32_t next_node_base, Node* __restrict__ nodes,
                           Task* __restrict__ next_tasks) {
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num_tasks) {
        return;
    }
    Task const     t = tasks[i];
    Decision const d = decisions[i];

    Node node;
    for (int a = 0; a < 3; ++a) {
        node.lo[a] = t.lo[a];
        node.hi[a] = t.hi[a];
    }

    if (Kind_split == d.kind) {
        uint32_t const rank  = split_rank[i];
        uint32_t const child = next_node_base + 2 * rank;
        node.data            = child;
        node.meta            = d.axis;

        Task c0;
        c0.begin = next_offset[i];
        c0.count = d.n0;
        c0.node  = child;
        c0.depth = t.depth + 1;
        Task c1;
        c1.begin = next_offset[i] + d.n0;
        c1.count = d.n1;
        c1.node  = child + 1;
        c1.depth = t.depth + 1;
        for (int a = 0; a < 3; ++a) {
            c0.lo[a] = d.lo0[a];
            c0.hi[a] = d.hi0[a];
            c1.lo[a] = d.lo1[a];
            c1.hi[a] = d.hi1[a];
        }
        next_tasks[2 * rank]     = c0;
        next_tasks[2 * rank + 1] = c1;
    } else {
        node.data = leaf_base + leaf_offset[i];
        node.meta = t.count << 8;
    }
    nodes[t.node] = node;
}

__device__ __forceinline__ void side_flags(Ref const& r, Decision const& d, bool& go0, bool& go1) {
    bool const behind_hi = r.hi[d.axis] < d.d;
    go0                  = behind_hi || (d.spatial && r.lo[d.axis] < d.d);
    go1                  = !behind_hi;
}

__global__ void __launch_bounds__(256)
count_sides(Chunk_rec const* __restrict__ chunks, Ref const* __restrict__ refs,
            Decision const* __restrict__ decisions, uint32_t* __restrict__ cnt0,
            uint32_t* __restrict__ cnt1) {
    __shared__ uint32_t s0[4];
    __shared__ uint32_t s1[4];

    Chunk_rec const ch = chunks[blockIdx.x];
    Decision const& d  = decisions[ch.task];

    uint32_t c0 = 0, c1 = 0;
    if (Kind_split == d.kind) {
        for (uint32_t i = ch.begin + threadIdx.x; i < ch.end; i += blockDim.x) {
            bool go0, go1;
            side_flags(refs[i], d, go0, go1);
            c0 += go0;
            c1 += go1;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        c0 += __shfl_xor(c0, off);
        c1 += __shfl_xor(c1, off);
    }
    if (0 == (threadIdx.x & 63)) {
        s0[threadIdx.x >> 6] = c0;
        s1[threadIdx.x >> 6] = c1;
    }
    __syncthreads();
    if (0 == threadIdx.x) {
        uint32_t t0 = 0, t1 = 0;
        for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) {
            t0 += s0[w];
            t1 += s1[w];
        }
        cnt0[blockIdx.x] = t0;
        cnt1[blockIdx.x] = t1;
    }
}

// Exclusive prefix over the block of per-thread values (blockDim 256).
__device__ uint32_t block_exclusive(uint32_t v, uint32_t* lds) {
    uint32_t const lane = threadIdx.x & 63;
    uint32_t const wave = threadIdx.x >> 6;
    uint32_t       inc  = v;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t const o = __shfl_up(inc, off);
        if (lane >= uint32_t(off)) {
            inc += o;
        }
    }
    if (63 == lane) {
        lds[wave] = inc;
    }
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t w = 0; w < wave; ++w) {
        base += lds[w];
    }
    __syncthreads();
    return base + inc - v;
}

// Stable partition of split tasks' references into the two children
// (Split_candidate::distribute, scene_bvh_split_candidate.inl:125-151) and
// in-order copy of leaf references (Kernel::assign, scene_bvh_builder_base.cpp:181-189).
__global__ void __launch_bounds__(256)
move_refs(Chunk_rec const* __restrict__ chunks, Ref const* __restrict__ refs,
          Decision const* __restrict__ decisions, Task const* __restrict__ tasks,
          uint32_t const* __restrict__ all_chunk_offset, uint32_t const* __restrict__ pre0,
          uint32_t const* __restrict__ pre1, uint32_t const* __restrict__ next_offset,
          uint32_t const* __restrict__ leaf_offset, uint32_t leaf_base, Ref* __restrict__ out,
          uint32_t* __restrict__ leaf_refs, uint32_t* __restrict__ error) {
    __shared__ uint32_t lds[8];

    Chunk_rec const ch = chunks[blockIdx.x];
    Decision const  d  = decisions[ch.task];
    Task const      t  = tasks[ch.task];

    constexpr uint32_t Per_thread = Chunk / 256;
    uint32_t const     first      = ch.begin + threadIdx.x * Per_thread;

    if (Kind_leaf == d.kind) {
        uint32_t const dst = leaf_base + leaf_offset[ch.task] + (ch.begin - t.begin);
        for (uint32_t k = 0; k < Per_thread; ++k) {
            uint32_t const i = first + k;
            if (i < ch.end) {
                leaf_refs[dst + (i - ch.begin)] = refs[i].prim;
            }
        }
        return;
    }
    if (Kind_split != d.kind) {
        return;
    }

    Ref  local[Per_thread];
    bool g0[Per_thread], g1[Per_thread];
    uint32_t c0 = 0, c1 = 0;
    for (uint32_t k = 0; k < Per_thread; ++k) {
        uint32_t const i = first + k;
        g0[k] = g1[k] = false;
        if (i < ch.end) {
            local[k] = refs[i];
            side_flags(local[k], d, g0[k], g1[k]);
        }
        c0 += g0[k];
        c1 += g1[k];
    }

    uint32_t const first_chunk = all_chunk_offset[ch.task];
    uint32_t const end0        = next_offset[ch.task] + d.n0;
    uint32_t const end1        = end0 + d.n1;
    uint32_t       o0 = next_offset[ch.task] + (pre0[blockIdx.x] - pre0[first_chunk]) +
                  block_exclusive(c0, lds);
    uint32_t o1 = end0 + (pre1[blockIdx.x] - pre1[first_chunk]) + block_exclusive(c1, lds + 4);

    // The side counts were taken with the same comparisons, so o0 < end0 and
    // o1 < end1 always; the guards keep a broken invariant from writing
    // outside the children's segments.
    for (uint32_t k = 0; k < Per_thread; ++k) {
        Ref const& r        = local[k];
        bool const straddle = g0[k] && g1[k];
        if (g0[k]) {
            Ref x = r;
            if (straddle) {
                x.hi[d.axis] = stdmin(d.d, x.hi[d.axis]);
            }
            if (o0 < end0) {
                out[o0] = x;
            } else {
                atomicOr(error, 4u);
            }
            ++o0;
        }
        if (g1[k]) {
            Ref y = r;
            if (straddle) {
                y.lo[d.axis] = stdmax(d.d, y.lo[d.axis]);
            }
            if (o1 < end1) {
                out[o1] = y;
            } else {
                atomicOr(error, 4u);
            }
            ++o1;
        }
    }
}

// ---- depth-first serialization ----------------------------------------------

__global__ void subtree_sizes(Node const* __restrict__ nodes, uint32_t begin, uint32_t end,
                              uint32_t* __restrict__ sub_inner, uint32_t* __restrict__ sub_slots) {
    uint32_t const x = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= end) {
        return;
    }
    Node const     n   = nodes[x];
    uint32_t const num = (n.meta >> 8) & 0xFF;
    if (num) {
        sub_inner[x] = 0;
        sub_slots[x] = num;
    } else {
        sub_inner[x] = 1 + sub_inner[n.data] + sub_inner[n.data + 1];
        sub_slots[x] = sub_slots[n.data] + sub_slots[n.data + 1];
    }
}

// dest: depth-first index; pair: where this inner node's children go;
// slot: first tree slot of this sub-tree (serialize(),
// triangle_bvh_builder_sah.cpp:129-164).
__global__ void place(Node const* __restrict__ nodes, uint32_t begin, uint32_t end,
                      uint32_t const* __restrict__ sub_inner,
                      uint32_t const* __restrict__ sub_slots, uint32_t* __restrict__ dest,
                      uint32_t* __restrict__ pair, uint32_t* __restrict__ slot) {
    uint32_t const x = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= end) {
        return;
    }
    Node const     n   = nodes[x];
    uint32_t const num = (n.meta >> 8) & 0xFF;
    if (num) {
        return;
    }
    uint32_t const c0 = n.data;
    uint32_t const c1 = n.data + 1;
    dest[c0]          = pair[x];
    dest[c1]          = pair[x] + 1;
    pair[c0]          = pair[x] + 2;
    pair[c1]          = pair[x] + 2 + 2 * sub_inner[c0];
    slot[c0]          = slot[x];
    slot[c1]          = slot[x] + sub_slots[c0];
}

__global__ void scatter(Node const* __restrict__ nodes, uint32_t num_nodes,
                        uint32_t const* __restrict__ dest, uint32_t const* __restrict__ pair,
                        uint32_t const* __restrict__ slot, uint32_t const* __restrict__ leaf_refs,
                        uint32_t num_slots, Node* __restrict__ out_nodes,
                        uint32_t* __restrict__ out_slots, uint32_t* __restrict__ error) {
    uint32_t const x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= num_nodes) {
        return;
    }
    Node           n   = nodes[x];
    uint32_t const num = (n.meta >> 8) & 0xFF;
    if (dest[x] >= num_nodes || (num && (uint64_t(slot[x]) + num > num_slots ||
                                         uint64_t(n.data) + num > num_slots))) {
        atomicOr(error, 8u);
        return;
    }
    if (num) {
        uint32_t const s = slot[x];
        for (uint32_t k = 0; k < num; ++k) {
            out_slots[s + k] = leaf_refs[n.data + k];
        }
        n.data = s;
    } else {
        n.data = pair[x];
    }
    out_nodes[dest[x]] = n;
}

__global__ void set_root_place(uint32_t* dest, uint32_t* pair, uint32_t* slot) {
    dest[0] = 0;
    pair[0] = 1;
    slot[0] = 0;
}

// ---- host side --------------------------------------------------------------

template <typename T>
struct Buffer {
    T*     p   = nullptr;
    size_t cap = 0;

    ~Buffer() { (void)hipFree(p); }

    hipError_t reserve(size_t n) {
        if (n <= cap) {
            return hipSuccess;
        }
        (void)hipFree(p);
        p   = nullptr;
        cap = 0;
        size_t const want = n + n / 4 + 64;
        hipError_t   e    = hipMalloc(&p, want * sizeof(T));
        if (e == hipSuccess) {
            cap = want;
        }
        return e;
    }

    // keeps the contents
    hipError_t grow(size_t n, size_t keep, hipStream_t s) {
        if (n <= cap) {
            return hipSuccess;
        }
        T*           q    = nullptr;
        size_t const want = n + n / 2 + 64;
        hipError_t   e    = hipMalloc(&q, want * sizeof(T));
        if (e != hipSuccess) {
            return e;
        }
        if (keep) {
            e = hipMemcpyAsync(q, p, keep * sizeof(T), hipMemcpyDeviceToDevice, s);
            if (e == hipSuccess) {
                e = hipStreamSynchronize(s);
            }
        }
        (void)hipFree(p);
        p   = q;
        cap = want;
        return e;
    }
};

struct Scanner {
    void*  temp       = nullptr;
    size_t temp_bytes = 0;

    ~Scanner() { (void)hipFree(temp); }

    hipError_t exclusive(uint32_t const* in, uint32_t* out, uint32_t n, hipStream_t s) {
        if (0 == n) {
            return hipSuccess;
        }
        size_t     need = 0;
        hipError_t e    = hipcub::DeviceScan::ExclusiveSum(nullptr, need, in, out, n, s);
        if (e != hipSuccess) {
            return e;
        }
        if (need > temp_bytes) {
            (void)hipFree(temp);
            temp       = nullptr;
            temp_bytes = 0;
            e          = hipMalloc(&temp, need);
            if (e != hipSuccess) {
                return e;
            }
            temp_bytes = need;
        }
        return hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, in, out, n, s);
    }
};

__global__ void gather_totals(uint32_t const* const* offsets, uint32_t const* const* counts,
                              uint32_t n, uint32_t num_arrays, uint32_t* totals) {
    uint32_t const k = threadIdx.x;
    if (k < num_arrays) {
        totals[k] = n ? offsets[k][n - 1] + counts[k][n - 1] : 0;
    }
}

uint32_t blocks_for(uint64_t n, uint32_t block) {
    uint64_t const b = (n + block - 1) / block;
    return uint32_t(b > 0 ? b : 1);
}

}  // namespace

#define MI_TRY(expr)                                                                  \
    do {                                                                              \
        hipError_t const e_ = (expr);                                                 \
        if (e_ != hipSuccess) {                                                       \
            return fail(MI_ERR_HIP, "bvh build: %s: %s", #expr, hipGetErrorString(e_)); \
        }                                                                             \
    } while (0)

#define MI_LAUNCHED()                                                                   \
    do {                                                                                \
        hipError_t const e_ = hipGetLastError();                                        \
        if (e_ != hipSuccess) {                                                         \
            return fail(MI_ERR_HIP, "bvh build launch: %s", hipGetErrorString(e_));     \
        }                                                                               \
    } while (0)

int32_t build_bvh(float const* positions, uint32_t num_vertices, uint32_t position_stride,
                  uint32_t const* indices, uint32_t num_triangles, mi_bvh_settings const& settings,
                  hipStream_t stream, DeviceBvh& out) {
    if (settings.sweep_threshold > 1024 || settings.num_slices > 64) {
        return fail(MI_ERR_ARGS, "mi_bvh_build: sweep_threshold <= 1024 and num_slices <= 64");
    }

    auto const start = std::chrono::steady_clock::now();

    Settings s;
    s.num_slices      = settings.num_slices;
    s.sweep_threshold = settings.sweep_threshold;
    s.max_primitives  = settings.max_primitives;
    // Builder_base::split, scene_bvh_builder_base.cpp:271-276
    s.spatial_split_threshold = uint32_t(std::lrint(std::log2(float(num_triangles)) / 2.f));

    hipStream_t const st = stream;

    Buffer<Ref>       refs_a, refs_b;
    Buffer<Task>      tasks_a, tasks_b;
    Buffer<Node>      nodes;
    Buffer<uint32_t>  leaf_refs;
    Buffer<uint32_t>  kind, cand_count, cand_offset, grid_chunks, grid_chunk_offset, all_chunks,
        all_chunk_offset;
    Buffer<uint32_t>  leaf_count, leaf_offset, split_flag, split_rank, child_refs, next_offset;
    Buffer<uint32_t>  cnt0, cnt1, pre0, pre1;
    Buffer<Cand>      cands;
    Buffer<Acc>       accs;
    Buffer<Chunk_rec> grid_list, all_list;
    Buffer<Decision>  decisions;
    Buffer<uint32_t>  scratch;  // root keys, error flag, totals
    Buffer<uint32_t const*> ptrs;
    Scanner           scan;

    MI_TRY(refs_a.reserve(num_triangles));
    MI_TRY(tasks_a.reserve(1));
    MI_TRY(nodes.reserve(std::max<size_t>(64, size_t(num_triangles) * 3 / std::max(1u, s.max_primitives))));
    MI_TRY(leaf_refs.reserve(size_t(num_triangles) + num_triangles / 4));
    MI_TRY(scratch.reserve(16));
    MI_TRY(ptrs.reserve(16));

    uint32_t* const root_keys = scratch.p;       // 6 words
    uint32_t* const error     = scratch.p + 6;   // 1 word
    uint32_t* const totals    = scratch.p + 8;   // up to 8 words

    {
        uint32_t init[8];
        for (int a = 0; a < 3; ++a) {
            float const big = Flt_max, small = -Flt_max;
            uint32_t    b, c;
            std::memcpy(&b, &big, 4);
            std::memcpy(&c, &small, 4);
            init[a]     = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
            init[3 + a] = (c & 0x80000000u) ? ~c : (c | 0x80000000u);
        }
        init[6] = 0;
        init[7] = 0;
        MI_TRY(hipMemcpyAsync(scratch.p, init, sizeof(init), hipMemcpyHostToDevice, st));
        MI_TRY(hipStreamSynchronize(st));
    }

    hipLaunchKernelGGL(make_refs, dim3(blocks_for(num_triangles, 256)), dim3(256), 0, st,
                       positions, position_stride, num_vertices, indices, num_triangles, refs_a.p,
                       root_keys, error);
    MI_LAUNCHED();
    hipLaunchKernelGGL(init_root, dim3(1), dim3(1), 0, st, root_keys, num_triangles, tasks_a.p);
    MI_LAUNCHED();

    Buffer<Ref>*  refs_cur  = &refs_a;
    Buffer<Ref>*  refs_next = &refs_b;
    Buffer<Task>* tasks_cur = &tasks_a;
    Buffer<Task>* tasks_nxt = &tasks_b;

    uint32_t              num_tasks  = 1;
    uint32_t              node_count = 1;
    uint32_t              leaf_total = 0;
    std::vector<uint32_t> level_start = {0};

    auto totals_of = [&](std::initializer_list<std::pair<uint32_t const*, uint32_t const*>> arrays,
                         uint32_t n, uint32_t* host) -> hipError_t {
        uint32_t const* h[16];
        uint32_t        k = 0;
        for (auto const& a : arrays) {
            h[k]                  = a.first;
            h[arrays.size() + k]  = a.second;
            ++k;
        }
        hipError_t e = hipMemcpyAsync(ptrs.p, h, 2 * k * sizeof(uint32_t const*),
                                      hipMemcpyHostToDevice, st);
        if (e != hipSuccess) {
            return e;
        }
        hipLaunchKernelGGL(gather_totals, dim3(1), dim3(64), 0, st, ptrs.p, ptrs.p + k, n, k,
                           totals);
        e = hipMemcpyAsync(host, totals, (k + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) {
            return e;
        }
        return hipStreamSynchronize(st);
    };

    while (num_tasks > 0) {
        uint32_t const n = num_tasks;

        MI_TRY(kind.reserve(n));
        MI_TRY(cand_count.reserve(n));
        MI_TRY(cand_offset.reserve(n));
        MI_TRY(grid_chunks.reserve(n));
        MI_TRY(grid_chunk_offset.reserve(n));
        MI_TRY(all_chunks.reserve(n));
        MI_TRY(all_chunk_offset.reserve(n));
        MI_TRY(decisions.reserve(n));
        MI_TRY(leaf_count.reserve(n));
        MI_TRY(leaf_offset.reserve(n));
        MI_TRY(split_flag.reserve(n));
        MI_TRY(split_rank.reserve(n));
        MI_TRY(child_refs.reserve(n));
        MI_TRY(next_offset.reserve(n));

        hipLaunchKernelGGL(plan, dim3(blocks_for(n, 256)), dim3(256), 0, st, tasks_cur->p, n, s,
                           kind.p, cand_count.p, grid_chunks.p, all_chunks.p);
        MI_LAUNCHED();
        MI_TRY(scan.exclusive(cand_count.p, cand_offset.p, n, st));
        MI_TRY(scan.exclusive(grid_chunks.p, grid_chunk_offset.p, n, st));
        MI_TRY(scan.exclusive(all_chunks.p, all_chunk_offset.p, n, st));

        uint32_t t1[4] = {0, 0, 0, 0};
        MI_TRY(totals_of({{cand_offset.p, cand_count.p},
                          {grid_chunk_offset.p, grid_chunks.p},
                          {all_chunk_offset.p, all_chunks.p}},
                         n, t1));
        uint32_t const num_cands       = t1[0];
        uint32_t const num_grid_chunks = t1[1];
        uint32_t const num_all_chunks  = t1[2];

        MI_TRY(cands.reserve(std::max(1u, num_cands)));
        MI_TRY(accs.reserve(std::max(1u, num_cands)));
        MI_TRY(grid_list.reserve(std::max(1u, num_grid_chunks)));
        MI_TRY(all_list.reserve(std::max(1u, num_all_chunks)));
        MI_TRY(cnt0.reserve(std::max(1u, num_all_chunks)));
        MI_TRY(cnt1.reserve(std::max(1u, num_all_chunks)));
        MI_TRY(pre0.reserve(std::max(1u, num_all_chunks)));
        MI_TRY(pre1.reserve(std::max(1u, num_all_chunks)));

        hipLaunchKernelGGL(make_candidates, dim3(blocks_for(n, 64)), dim3(64), 0, st, tasks_cur->p,
                           n, s, kind.p, cand_offset.p, cands.p, accs.p);
        MI_LAUNCHED();
        hipLaunchKernelGGL(make_chunks, dim3(blocks_for(n, 256)), dim3(256), 0, st, tasks_cur->p,
                           n, kind.p, grid_chunk_offset.p, all_chunk_offset.p, grid_list.p,
                           all_list.p);
        MI_LAUNCHED();

        if (num_grid_chunks) {
            hipLaunchKernelGGL(evaluate_grid, dim3(num_grid_chunks), dim3(256), 0, st, grid_list.p,
                               refs_cur->p, cand_offset.p, cand_count.p, cands.p, accs.p);
            MI_LAUNCHED();
        }
        hipLaunchKernelGGL(select_grid, dim3(n), dim3(128), 0, st, tasks_cur->p, kind.p,
                           cand_offset.p, cand_count.p, cands.p, accs.p, decisions.p);
        MI_LAUNCHED();
        hipLaunchKernelGGL(select_sweep, dim3(n), dim3(256),
                           size_t(std::max(1u, s.sweep_threshold)) * sizeof(Ref), st, tasks_cur->p,
                           kind.p, refs_cur->p, decisions.p);
        MI_LAUNCHED();
        hipLaunchKernelGGL(outcome_counts, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                           tasks_cur->p, n, kind.p, decisions.p, leaf_count.p, split_flag.p,
                           child_refs.p, error);
        MI_LAUNCHED();
        MI_TRY(scan.exclusive(leaf_count.p, leaf_offset.p, n, st));
        MI_TRY(scan.exclusive(split_flag.p, split_rank.p, n, st));
        MI_TRY(scan.exclusive(child_refs.p, next_offset.p, n, st));

        uint32_t t2[4] = {0, 0, 0, 0};
        MI_TRY(totals_of({{leaf_offset.p, leaf_count.p},
                          {split_rank.p, split_flag.p},
                          {next_offset.p, child_refs.p}},
                         n, t2));
        uint32_t err = 0;
        MI_TRY(hipMemcpyAsync(&err, error, 4, hipMemcpyDeviceToHost, st));
        MI_TRY(hipStreamSynchronize(st));
        if (err & 2u) {
            return fail(MI_ERR_ARGS, "mi_bvh_build: vertex index out of range");
        }
        if (err & 4u) {
            return fail(MI_ERR_HIP, "mi_bvh_build: internal error (side counts)");
        }
        if (err) {
            return fail(MI_ERR_UNSPLITTABLE,
                        "mi_bvh_build: a node of more than 255 references cannot be split "
                        "(level %zu)",
                        level_start.size() - 1);
        }

        uint32_t const level_leaf_refs = t2[0];
        uint32_t const num_splits      = t2[1];
        uint32_t const num_next_refs   = t2[2];

        MI_TRY(nodes.grow(size_t(node_count) + 2 * size_t(num_splits), node_count, st));
        MI_TRY(leaf_refs.grow(size_t(leaf_total) + level_leaf_refs, leaf_total, st));
        MI_TRY(tasks_nxt->reserve(std::max(1u, 2 * num_splits)));
        MI_TRY(refs_next->reserve(std::max(1u, num_next_refs)));

        hipLaunchKernelGGL(emit_nodes, dim3(blocks_for(n, 256)), dim3(256), 0, st, tasks_cur->p, n,
                           decisions.p, leaf_offset.p, split_rank.p, next_offset.p, leaf_total,
                           node_count, nodes.p, tasks_nxt->p);
        MI_LAUNCHED();

        if (num_all_chunks) {
            hipLaunchKernelGGL(count_sides, dim3(num_all_chunks), dim3(256), 0, st, all_list.p,
                               refs_cur->p, decisions.p, cnt0.p, cnt1.p);
            MI_LAUNCHED();
            MI_TRY(scan.exclusive(cnt0.p, pre0.p, num_all_chunks, st));
            MI_TRY(scan.exclusive(cnt1.p, pre1.p, num_all_chunks, st));
            hipLaunchKernelGGL(move_refs, dim3(num_all_chunks), dim3(256), 0, st, all_list.p,
                               refs_cur->p, decisions.p, tasks_cur->p, all_chunk_offset.p, pre0.p,
                               pre1.p, next_offset.p, leaf_offset.p, leaf_total, refs_next->p,
                               leaf_refs.p, error);
            MI_LAUNCHED();
        }

        level_start.push_back(node_count);
        node_count += 2 * num_splits;
        leaf_total += level_leaf_refs;
        num_tasks = 2 * num_splits;
        std::swap(refs_cur, refs_next);
        std::swap(tasks_cur, tasks_nxt);

        if (level_start.size() > Reference_stack_size + 2) {
            return fail(MI_ERR_STACK, "mi_bvh_build: tree deeper than %u levels",
                        Reference_stack_size);
        }
    }

    // level l holds build nodes [level_start[l], level_start[l + 1])
    uint32_t const num_levels = uint32_t(level_start.size() - 1);

    // depth-first layout
    Buffer<uint32_t> sub_inner, sub_slots, dest, pair, slot;
    MI_TRY(sub_inner.reserve(node_count));
    MI_TRY(sub_slots.reserve(node_count));
    MI_TRY(dest.reserve(node_count));
    MI_TRY(pair.reserve(node_count));
    MI_TRY(slot.reserve(node_count));

    for (uint32_t l = num_levels; l-- > 0;) {
        uint32_t const b = level_start[l], e = level_start[l + 1];
        hipLaunchKernelGGL(subtree_sizes, dim3(blocks_for(e - b, 256)), dim3(256), 0, st, nodes.p,
                           b, e, sub_inner.p, sub_slots.p);
        MI_LAUNCHED();
    }
    hipLaunchKernelGGL(set_root_place, dim3(1), dim3(1), 0, st, dest.p, pair.p, slot.p);
    MI_LAUNCHED();
    for (uint32_t l = 0; l < num_levels; ++l) {
        uint32_t const b = level_start[l], e = level_start[l + 1];
        hipLaunchKernelGGL(place, dim3(blocks_for(e - b, 256)), dim3(256), 0, st, nodes.p, b, e,
                           sub_inner.p, sub_slots.p, dest.p, pair.p, slot.p);
        MI_LAUNCHED();
    }

    DeviceBvh bvh;
    bvh.num_nodes     = node_count;
    bvh.num_slots     = leaf_total;
    bvh.num_triangles = num_triangles;
    bvh.max_depth     = num_levels - 1;

    hipError_t e = hipMalloc(&bvh.nodes_alloc, size_t(node_count + 1) * 32);
    if (e == hipSuccess) {
        e = hipMalloc(&bvh.tris, size_t(std::max(1u, leaf_total)) * 48);
    }
    if (e == hipSuccess) {
        e = hipMalloc(&bvh.slots, size_t(std::max(1u, leaf_total)) * 4);
    }
    if (e != hipSuccess) {
        bvh.release();
        return fail(MI_ERR_HIP, "bvh build: %s", hipGetErrorString(e));
    }
    bvh.nodes = reinterpret_cast<float4*>(static_cast<char*>(bvh.nodes_alloc) + 32);

    hipLaunchKernelGGL(scatter, dim3(blocks_for(node_count, 256)), dim3(256), 0, st, nodes.p,
                       node_count, dest.p, pair.p, slot.p, leaf_refs.p, leaf_total,
                       reinterpret_cast<Node*>(bvh.nodes), bvh.slots, error);
    if (hipError_t const le = hipGetLastError(); le != hipSuccess) {
        bvh.release();
        return fail(MI_ERR_HIP, "bvh build launch: %s", hipGetErrorString(le));
    }
    if (int32_t const r = fill_triangles(bvh, positions, position_stride, indices, st); r) {
        bvh.release();
        return r;
    }

    float    root[8];
    uint32_t err = 0;
    e            = hipMemcpyAsync(root, bvh.nodes, 32, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) {
        e = hipMemcpyAsync(&err, error, 4, hipMemcpyDeviceToHost, st);
    }
    if (e == hipSuccess) {
        e = hipStreamSynchronize(st);
    }
    if (e != hipSuccess) {
        bvh.release();
        return fail(MI_ERR_HIP, "bvh build: %s", hipGetErrorString(e));
    }
    if (err) {
        bvh.release();
        return fail(MI_ERR_HIP, "mi_bvh_build: internal error (layout %u)", err);
    }
    for (int a = 0; a < 3; ++a) {
        bvh.aabb_min[a] = root[a];
        bvh.aabb_max[a] = root[4 + a];
    }
    bvh.build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();

    out = bvh;
    return MI_OK;
}

}  // namespace mi
