// Light trees built on the device: sprout's Tree_builder
// (core/scene/light/light_tree_builder.cpp:436-760) for the scene tree over
// a scene's lights and for the Primitive_trees of its emissive parts, so a
// scene whose emitters move rebuilds its trees without a host round trip
// (Scene::compile rebuilds the light tree every frame, scene.cpp:216-267).
//
// The host builder (light_tree.cpp) recurses depth first; this one splits a
// whole level of nodes at a time.  It gives the same tree bit for bit:
//   * the arithmetic is light_build.hpp's, shared with the host builder;
//   * a candidate's sums and cone merges run over the node's lights in the
//     order the host's recursion leaves them (one thread per candidate);
//   * the cheapest candidate is the first of equal costs (one thread scans
//     a node's candidates in order);
//   * the partition is libstdc++'s std::partition for bidirectional
//     iterators (swap the first light that fails from the left with the last
//     that passes from the right), one thread per node;
//   * nodes are numbered as the recursion numbers them: the children of the
//     inner node with pre-order rank r among inner nodes are 2r + 1, 2r + 2;
//   * a light's order is its final position (leaves are assigned left to
//     right), a node's variance runs over its final order.
// A scene-tree node whose every candidate leaves one side empty (the
// reference recurses forever there; the host builder falls back to a
// median split) makes the build report it, and the caller builds on the
// host instead.
//
// Several trees ("roots") over disjoint light ranges build in one pass:
// every part's Primitive_tree at once.

#include "light_build.hip.hpp"

#include "light_build.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

namespace mi {

namespace {

constexpr uint32_t Null = 0xFFFFFFFFu;

// A node of the level-by-level build, in creation order (level by level).
struct Tnode {
    lb::Box  bounds;
    lb::Cone cone;
    float    power;
    uint32_t begin, end;  // light positions (global)
    uint32_t root;
    uint32_t left;        // temp id of the first child, or Null (a leaf)
    uint32_t mid;         // first position of the second child
    uint32_t two_sided;
    uint32_t inner_count;  // inner nodes in the subtree
    uint32_t rank;         // pre-order rank among the root's inner nodes
    uint32_t id;           // final node index within the root's tree
};

struct Forest {
    lb::Light_set    set;
    uint32_t const*  first;      // R + 1 light offsets of the roots
    float const*     part_cones; // 4 per root (parts)
    uint32_t         num_roots;
    uint32_t         part;       // Primitive_trees (else the scene tree)
    uint32_t*        mapping;    // positions -> global light
    Tnode*           nodes;
    uint32_t*        counters;   // [0] nodes made, [1] candidates, [2] failed
    uint32_t*        cand_node;  // per candidate: its frontier entry
    uint32_t*        cand_k;     // per candidate: its index within the node
    float*           cand_cost;
    uint32_t*        cand_start; // per frontier entry
    uint32_t*        cand_count;
};

__device__ uint32_t leaf_size(Forest const& f, uint32_t root) {
    if (!f.part) {
        return 4;
    }
    uint32_t const n = f.first[root + 1] - f.first[root];
    return n / 64 > 4 ? n / 64 : 4;
}

__device__ uint32_t sweep_threshold(Forest const& f) {
    return f.part ? lb::Part_sweep_threshold : lb::Scene_sweep_threshold;
}

// Root r: its bounds, cone, power, sidedness (build_light_tree /
// build_part_tree's first loop).
__global__ void init_roots(Forest f) {
    uint32_t const r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= f.num_roots) {
        return;
    }
    uint32_t const b = f.first[r];
    uint32_t const e = f.first[r + 1];
    Tnode          t;
    t.begin = b;
    t.end   = e;
    t.root  = r;
    t.left  = Null;
    t.mid   = 0;
    t.power = 0.f;
    t.two_sided   = 0;
    t.inner_count = 0;
    t.rank        = 0;
    t.id          = 0;
    for (uint32_t i = b; i < e; ++i) {
        f.mapping[i] = i;
    }
    if (f.part) {
        for (uint32_t l = b; l < e; ++l) {
            if (f.set.powers[l] > 0.f) {
                t.bounds.merge(f.set.lo + 3 * size_t(l), f.set.hi + 3 * size_t(l));
                t.power += f.set.powers[l];
            }
        }
        for (int k = 0; k < 4; ++k) {
            t.cone.v[k] = f.part_cones[4 * size_t(r) + k];
        }
    } else {
        bool ts = false;
        for (uint32_t l = b; l < e; ++l) {
            t.bounds.merge(f.set.lo + 3 * size_t(l), f.set.hi + 3 * size_t(l));
            t.cone = lb::merge(t.cone, f.set.cone(l));
            ts     = ts || f.set.two(l);
            t.power += f.set.powers[l];
        }
        t.two_sided = ts ? 1u : 0u;
    }
    f.nodes[r] = t;
}

// Per node of the level: a leaf, or how many split candidates it has.
__global__ void count_candidates(Forest f, uint32_t level_begin, uint32_t level_size) {
    uint32_t const a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= level_size) {
        return;
    }
    Tnode const&   t   = f.nodes[level_begin + a];
    uint32_t const len = t.end - t.begin;
    uint32_t       n   = 0;
    if (len > leaf_size(f, t.root)) {
        n = len < sweep_threshold(f) ? 3 * len : lb::grid_candidates(t.bounds);
    }
    f.cand_count[a] = n;
}

// Exclusive scan of the level's candidate counts (one block).
__global__ void __launch_bounds__(1024) scan_candidates(Forest f, uint32_t level_size) {
    __shared__ uint32_t part[1024];
    uint32_t const      tid   = threadIdx.x;
    uint32_t const      chunk = (level_size + 1023) / 1024;
    uint32_t const      b     = tid * chunk;
    uint32_t const      e     = min(level_size, b + chunk);
    uint32_t            sum   = 0;
    for (uint32_t i = b; i < e; ++i) {
        sum += f.cand_count[i];
    }
    part[tid] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        uint32_t const v = tid >= off ? part[tid - off] : 0u;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    uint32_t run = part[tid] - sum;
    for (uint32_t i = b; i < e; ++i) {
        f.cand_start[i] = run;
        run += f.cand_count[i];
    }
    if (1023 == tid) {
        f.counters[1] = part[1023];
    }
}

// Each candidate learns its node and index.
__global__ void expand_candidates(Forest f, uint32_t level_size) {
    uint32_t const a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= level_size) {
        return;
    }
    uint32_t const s = f.cand_start[a];
    for (uint32_t k = 0, n = f.cand_count[a]; k < n; ++k) {
        f.cand_node[s + k] = a;
        f.cand_k[s + k]    = k;
    }
}

__device__ void make_candidate(Forest const& f, Tnode const& t, uint32_t k, lb::Candidate& c) {
    c = lb::Candidate();
    if (t.end - t.begin < sweep_threshold(f)) {
        lb::sweep_candidate(f.mapping, t.begin, k, f.set, c);
    } else {
        lb::grid_candidate(t.bounds, k, c);
    }
}

// The cost of every candidate of the level: one thread walks the node's
// lights in their current order.
__global__ void evaluate_candidates(Forest f, uint32_t level_begin) {
    uint32_t const total  = f.counters[1];
    uint32_t const stride = gridDim.x * blockDim.x;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < total; c += stride) {
        Tnode const&  t = f.nodes[level_begin + f.cand_node[c]];
        lb::Candidate cand;
        make_candidate(f, t, f.cand_k[c], cand);
        cand.evaluate(t.begin, t.end, f.mapping, t.bounds, lb::cone_cost(t.cone.v[3]), f.set);
        f.cand_cost[c] = cand.cost;
    }
}

// Per splitting node: the first cheapest candidate, its sides, the
// partition (std::partition), and the two children.
__global__ void split_nodes(Forest f, uint32_t level_begin, uint32_t level_size) {
    uint32_t const a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= level_size) {
        return;
    }
    Tnode&         t = f.nodes[level_begin + a];
    uint32_t const n = f.cand_count[a];
    if (0 == n && t.end - t.begin <= leaf_size(f, t.root)) {
        return;  // a leaf
    }
    lb::Candidate sc;
    if (n > 0) {
        uint32_t const s        = f.cand_start[a];
        float          min_cost = f.cand_cost[s];
        uint32_t       best     = 0;
        for (uint32_t k = 1; k < n; ++k) {
            if (f.cand_cost[s + k] < min_cost) {
                best     = k;
                min_cost = f.cand_cost[s + k];
            }
        }
        make_candidate(f, t, best, sc);
        sc.evaluate(t.begin, t.end, f.mapping, t.bounds, lb::cone_cost(t.cone.v[3]), f.set);
    } else {
        sc.exhausted = true;  // no candidate: every light on the far side
        sc.d         = -3.40282347e+38f;
    }
    if (f.part && sc.exhausted) {
        return;  // Tree_builder::split of a part: an exhausted node is a leaf
    }
    // std::partition over [begin, end) by behind(box max)
    uint32_t* first = f.mapping + t.begin;
    uint32_t* last  = f.mapping + t.end;
    for (;;) {
        for (;;) {
            if (first == last) {
                goto done;
            }
            if (sc.behind(f.set.hi + 3 * size_t(*first))) {
                ++first;
            } else {
                break;
            }
        }
        --last;
        for (;;) {
            if (first == last) {
                goto done;
            }
            if (!sc.behind(f.set.hi + 3 * size_t(*last))) {
                --last;
            } else {
                break;
            }
        }
        uint32_t const x = *first;
        *first           = *last;
        *last            = x;
        ++first;
    }
done:
    uint32_t const split = uint32_t(first - f.mapping);
    if (split == t.begin || split == t.end) {
        // the host builder's median fallback: not built here
        atomicOr(&f.counters[2], 1u);
        return;
    }
    uint32_t const c = atomicAdd(&f.counters[0], 2u);
    Tnode          l, r;
    l.bounds = sc.box0, l.cone = sc.cone0, l.power = sc.power0;
    r.bounds = sc.box1, r.cone = sc.cone1, r.power = sc.power1;
    l.begin = t.begin, l.end = split;
    r.begin = split, r.end = t.end;
    l.root = r.root = t.root;
    l.left = r.left = Null;
    l.mid = r.mid = 0;
    l.two_sided = (!f.part && sc.two0) ? 1u : 0u;
    r.two_sided = (!f.part && sc.two1) ? 1u : 0u;
    l.inner_count = r.inner_count = 0;
    l.rank = r.rank = 0;
    l.id = r.id = 0;
    f.nodes[c]     = l;
    f.nodes[c + 1] = r;
    t.left         = c;
    t.mid          = split;
}

// Inner nodes per subtree, a level at a time from the deepest.
__global__ void count_inner(Forest f, uint32_t level_begin, uint32_t level_size) {
    uint32_t const a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= level_size) {
        return;
    }
    Tnode& t = f.nodes[level_begin + a];
    t.inner_count =
        Null == t.left ? 0u : 1u + f.nodes[t.left].inner_count + f.nodes[t.left + 1].inner_count;
}

// Pre-order ranks and node ids, a level at a time from the roots: the
// children of rank r are 2r + 1 and 2r + 2.
__global__ void number_nodes(Forest f, uint32_t level_begin, uint32_t level_size) {
    uint32_t const a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= level_size) {
        return;
    }
    Tnode const& t = f.nodes[level_begin + a];
    if (Null == t.left) {
        return;
    }
    Tnode& l = f.nodes[t.left];
    Tnode& r = f.nodes[t.left + 1];
    l.rank   = t.rank + 1;
    r.rank   = t.rank + 1 + l.inner_count;
    l.id     = 1 + 2 * t.rank;
    r.id     = 2 + 2 * t.rank;
}

// The serialized nodes (serialize, light_tree_builder.cpp:712-733), the
// light orders and the local mapping.
__global__ void write_nodes(Forest f, uint32_t num_temp, uint32_t const* node_base,
                            mi_light_node* out, uint32_t* middles) {
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num_temp) {
        return;
    }
    Tnode const&   t     = f.nodes[i];
    uint32_t const first = f.first[t.root];
    bool           two   = 0 != t.two_sided;
    if (Null == t.left && !f.part) {
        // a scene leaf: two-sided if any of its lights is
        two = false;
        for (uint32_t p = t.begin; p < t.end; ++p) {
            two = two || f.set.two(f.mapping[p]);
        }
    }
    mi_light_node d;
    float const   e[3] = {t.bounds.extent(0), t.bounds.extent(1), t.bounds.extent(2)};
    d.center[0]        = t.bounds.position(0);
    d.center[1]        = t.bounds.position(1);
    d.center[2]        = t.bounds.position(2);
    d.radius           = 0.5f * sqrtf(lb::dot3(e, e));
    for (int k = 0; k < 4; ++k) {
        d.cone[k] = t.cone.v[k];
    }
    d.power    = t.power;
    d.variance = lb::variance(f.mapping, t.begin, t.end, f.set);
    uint32_t const inner    = Null != t.left ? 1u : 0u;
    uint32_t const children = inner ? f.nodes[t.left].id : t.begin - first;
    d.bits       = inner | (two ? 2u : 0u) | (children << 2);
    d.num_lights = t.end - t.begin;
    uint32_t const at = node_base[t.root] + t.id;
    out[at]           = d;
    middles[at]       = inner ? t.mid - first : 0u;
}

__global__ void write_lights(Forest f, uint32_t num_lights, uint32_t const* light_root,
                             uint32_t* mapping_out, uint32_t* orders_out) {
    uint32_t const p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= num_lights) {
        return;
    }
    uint32_t const first = f.first[light_root[p]];
    uint32_t const l     = f.mapping[p];
    mapping_out[p]       = l - first;
    orders_out[l]        = p - first;
}

// Leaf distributions of the parts (light_tree_builder.cpp:760-795,
// distribution_1d.cpp:42-82): per part, over its nodes in id order, leaves
// of more than 4 lights get num + 1 cdf values (2 when their power is 0).
__global__ void part_cdfs(Forest f, uint32_t const* node_base, uint32_t const* cdf_base,
                          mi_light_node const* nodes, uint32_t const* num_nodes,
                          uint32_t* cdf_offset, float* cdfs, uint32_t* num_cdfs) {
    uint32_t const r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= f.num_roots) {
        return;
    }
    uint32_t const first = f.first[r];
    uint32_t       c     = cdf_base[r];
    for (uint32_t i = node_base[r], e = node_base[r] + num_nodes[r]; i < e; ++i) {
        mi_light_node const& nd = nodes[i];
        cdf_offset[i]           = Null;
        if ((nd.bits & 1u) || nd.num_lights <= 4) {
            continue;
        }
        uint32_t const b        = first + (nd.bits >> 2);
        uint32_t const num      = nd.num_lights;
        float          integral = 0.f;
        for (uint32_t t = 0; t < num; ++t) {
            integral += f.set.powers[f.mapping[b + t]];
        }
        cdf_offset[i] = c - cdf_base[r];
        if (0.f == integral) {
            cdfs[c++] = 1.f;
            cdfs[c++] = 1.f;
            continue;
        }
        float const ii = 1.f / integral;
        float       v  = 0.f;
        cdfs[c++]      = v;
        for (uint32_t t = 1; t < num; ++t) {
            v         = __builtin_fmaf(f.set.powers[f.mapping[b + t - 1]], ii, v);
            cdfs[c++] = v;
        }
        cdfs[c++] = 1.f;
    }
    num_cdfs[r] = c - cdf_base[r];
}

uint32_t blocks(uint64_t n, uint32_t b) {
    return uint32_t(std::max<uint64_t>(1, (n + b - 1) / b));
}

}  // namespace

void Light_forest_scratch::release() {
    (void)hipFree(mapping);
    (void)hipFree(nodes);
    (void)hipFree(counters);
    (void)hipFree(cand_node);
    (void)hipFree(cand_k);
    (void)hipFree(cand_cost);
    (void)hipFree(cand_start);
    (void)hipFree(cand_count);
    (void)hipFree(node_base);
    (void)hipFree(light_root);
    (void)hipFree(cdf_base);
    (void)hipFree(num_nodes);
    (void)hipFree(first);
    if (counters_host) {
        (void)hipHostFree(counters_host);
    }
    *this = Light_forest_scratch();
}

int32_t Light_forest_scratch::reserve(uint32_t lights, uint32_t roots) {
    if (!counters_host) {
        if (hipHostMalloc(&counters_host, 4 * sizeof(uint32_t)) != hipSuccess) {
            counters_host = nullptr;
            return MI_ERR_HIP;
        }
    }
    if (lights <= light_capacity && roots <= root_capacity) {
        return MI_OK;
    }
    uint32_t const nl = std::max(lights, light_capacity);
    uint32_t const nr = std::max(roots, root_capacity);
    Light_forest_scratch fresh;
    fresh.counters_host = counters_host;
    counters_host       = nullptr;
    release();
    *this            = fresh;
    size_t const cap = 4 * size_t(nl) + 64;  // candidates per level: 3 per light + grids
    hipError_t   e   = hipMalloc(&mapping, size_t(nl) * sizeof(uint32_t));
    if (!e) e = hipMalloc(&nodes, (2 * size_t(nl) + 1) * sizeof(Tnode));
    if (!e) e = hipMalloc(&counters, 4 * sizeof(uint32_t));
    if (!e) e = hipMalloc(&cand_node, cap * sizeof(uint32_t));
    if (!e) e = hipMalloc(&cand_k, cap * sizeof(uint32_t));
    if (!e) e = hipMalloc(&cand_cost, cap * sizeof(float));
    if (!e) e = hipMalloc(&cand_start, size_t(nl + 1) * sizeof(uint32_t));
    if (!e) e = hipMalloc(&cand_count, size_t(nl + 1) * sizeof(uint32_t));
    if (!e) e = hipMalloc(&node_base, size_t(nr + 1) * sizeof(uint32_t));
    if (!e) e = hipMalloc(&light_root, size_t(nl) * sizeof(uint32_t));
    if (!e) e = hipMalloc(&cdf_base, size_t(nr + 1) * sizeof(uint32_t));
    if (!e) e = hipMalloc(&num_nodes, size_t(nr + 1) * sizeof(uint32_t));
    if (!e) e = hipMalloc(&first, size_t(nr + 1) * sizeof(uint32_t));
    if (e) {
        release();
        return MI_ERR_HIP;
    }
    light_capacity = nl;
    root_capacity  = nr;
    return MI_OK;
}

int32_t build_light_forest(Light_forest const& in, Light_forest_scratch& s, Light_forest_out& out,
                           hipStream_t stream) {
    uint32_t const R = in.num_roots;
    uint32_t const N = in.num_lights;
    if (0 == R || 0 == N) {
        return MI_OK;
    }
    if (int32_t const st = s.reserve(N, R); st) {
        return st;
    }
    // per root: its light offsets, node and cdf bases (host layout: root r
    // owns nodes [2 first_r - r, + 2 n_r - 1) and cdfs [2 first_r, + 2 n_r))
    std::vector<uint32_t> node_base(R + 1), cdf_base(R + 1), light_root(N);
    for (uint32_t r = 0; r < R; ++r) {
        node_base[r] = 2 * in.first[r] - r;
        cdf_base[r]  = 2 * in.first[r];
        for (uint32_t l = in.first[r]; l < in.first[r + 1]; ++l) {
            light_root[l] = r;
        }
        if (in.first[r + 1] <= in.first[r]) {
            return MI_ERR_ARGS;  // every root holds a light
        }
    }
    node_base[R] = 2 * N - R;
    cdf_base[R]  = 2 * N;
    hipError_t e = hipMemcpyAsync(s.first, in.first, (R + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, stream);
    if (!e) e = hipMemcpyAsync(s.node_base, node_base.data(), (R + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, stream);
    if (!e) e = hipMemcpyAsync(s.cdf_base, cdf_base.data(), (R + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, stream);
    if (!e) e = hipMemcpyAsync(s.light_root, light_root.data(), N * sizeof(uint32_t), hipMemcpyHostToDevice, stream);
    if (!e) e = hipMemsetAsync(s.counters, 0, 4 * sizeof(uint32_t), stream);
    if (e) {
        return MI_ERR_HIP;
    }

    Forest f;
    f.set        = in.set;
    f.first      = s.first;
    f.part_cones = in.part_cones;
    f.num_roots  = R;
    f.part       = in.part ? 1u : 0u;
    f.mapping    = s.mapping;
    f.nodes      = static_cast<Tnode*>(s.nodes);
    f.counters   = s.counters;
    f.cand_node  = s.cand_node;
    f.cand_k     = s.cand_k;
    f.cand_cost  = s.cand_cost;
    f.cand_start = s.cand_start;
    f.cand_count = s.cand_count;

    hipLaunchKernelGGL(init_roots, dim3(blocks(R, 64)), dim3(64), 0, stream, f);
    // the roots are nodes 0 .. R-1; counters[0] counts the nodes made
    uint32_t const made0 = R;
    e = hipMemcpyAsync(s.counters, &made0, sizeof(uint32_t), hipMemcpyHostToDevice, stream);
    if (e) {
        return MI_ERR_HIP;
    }
    std::vector<uint32_t> level_start{0, R};
    for (;;) {
        uint32_t const lb0 = level_start[level_start.size() - 2];
        uint32_t const ls  = level_start.back() - lb0;
        if (0 == ls) {
            break;
        }
        hipLaunchKernelGGL(count_candidates, dim3(blocks(ls, 256)), dim3(256), 0, stream, f, lb0, ls);
        hipLaunchKernelGGL(scan_candidates, dim3(1), dim3(1024), 0, stream, f, ls);
        hipLaunchKernelGGL(expand_candidates, dim3(blocks(ls, 256)), dim3(256), 0, stream, f, ls);
        hipLaunchKernelGGL(evaluate_candidates, dim3(std::min<uint32_t>(blocks(4ull * N + 64, 64), 8192)),
                           dim3(64), 0, stream, f, lb0);
        hipLaunchKernelGGL(split_nodes, dim3(blocks(ls, 64)), dim3(64), 0, stream, f, lb0, ls);
        e = hipMemcpyAsync(s.counters_host, s.counters, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, stream);
        if (!e) e = hipStreamSynchronize(stream);
        if (e) {
            return MI_ERR_HIP;
        }
        if (s.counters_host[2]) {
            return MI_ERR_UNSPLITTABLE;
        }
        level_start.push_back(s.counters_host[0]);
    }
    uint32_t const num_temp   = level_start.back();
    size_t const   num_levels = level_start.size() - 1;
    for (size_t L = num_levels; L-- > 0;) {
        uint32_t const b = level_start[L], n = level_start[L + 1] - b;
        if (n) {
            hipLaunchKernelGGL(count_inner, dim3(blocks(n, 256)), dim3(256), 0, stream, f, b, n);
        }
    }
    for (size_t L = 0; L < num_levels; ++L) {
        uint32_t const b = level_start[L], n = level_start[L + 1] - b;
        if (n) {
            hipLaunchKernelGGL(number_nodes, dim3(blocks(n, 256)), dim3(256), 0, stream, f, b, n);
        }
    }
    hipLaunchKernelGGL(write_nodes, dim3(blocks(num_temp, 256)), dim3(256), 0, stream, f, num_temp,
                       s.node_base, out.nodes, out.middles);
    hipLaunchKernelGGL(write_lights, dim3(blocks(N, 256)), dim3(256), 0, stream, f, N, s.light_root,
                       out.mapping, out.orders);
    // nodes per root: the nodes the level loop made under it
    {
        std::vector<uint32_t> root_nodes(R, 1);
        // a root's tree has 2 * (inner nodes) + 1 nodes; read back the inner counts
        std::vector<Tnode> roots(R);
        e = hipMemcpyAsync(roots.data(), s.nodes, R * sizeof(Tnode), hipMemcpyDeviceToHost, stream);
        if (!e) e = hipStreamSynchronize(stream);
        if (e) {
            return MI_ERR_HIP;
        }
        for (uint32_t r = 0; r < R; ++r) {
            root_nodes[r] = 2 * roots[r].inner_count + 1;
        }
        if (out.num_nodes_host) {
            std::copy(root_nodes.begin(), root_nodes.end(), out.num_nodes_host);
        }
        if (in.part && out.cdf_offset) {
            e = hipMemcpyAsync(s.num_nodes, root_nodes.data(), R * sizeof(uint32_t), hipMemcpyHostToDevice, stream);
            if (e) {
                return MI_ERR_HIP;
            }
            hipLaunchKernelGGL(part_cdfs, dim3(blocks(R, 64)), dim3(64), 0, stream, f, s.node_base,
                               s.cdf_base, out.nodes, s.num_nodes, out.cdf_offset, out.cdfs,
                               out.num_cdfs);
        }
    }
    e = hipGetLastError();
    return e ? MI_ERR_HIP : MI_OK;
}

}  // namespace mi
