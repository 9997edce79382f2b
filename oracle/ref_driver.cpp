// Thin C driver around the reference's own triangle BVH, compiled from the
// reference sources where they lie (see ref.mk).  Nothing here restates the
// algorithm: it feeds arrays into
//   scene::shape::triangle::bvh::Builder_SAH::build   (triangle_bvh_builder_sah.cpp:21-67)
//   scene::shape::triangle::bvh::Tree::intersect      (triangle_bvh_tree.cpp:41-90)
//   scene::shape::triangle::bvh::Tree::intersect_p    (triangle_bvh_tree.cpp:145-185)
// and copies the results out.  Test infrastructure only: it is the checker
// for the oracle restatement and the "reference" CPU baseline.
//
// The triangle "part" field carries the input triangle id, so that the tree
// slot -> input triangle map can be read back through Tree::triangle_part
// (the part is only used for material lookup, never by the BVH).

#define private public
#include "scene/shape/triangle/bvh/triangle_bvh_tree.hpp"
#undef private

#include "base/math/aabb.inl"
#include "base/math/matrix4x4.inl"
#include "base/math/vector3.inl"
#include "scene/bvh/scene_bvh_builder.hpp"
#include "scene/bvh/scene_bvh_tree.inl"
#include "base/thread/thread_pool.hpp"
#include "scene/bvh/scene_bvh_node.inl"
#include "scene/shape/node_stack.inl"
#include "scene/shape/shape_vertex.hpp"
#include "scene/shape/triangle/bvh/triangle_bvh_builder_sah.hpp"
#include "scene/shape/triangle/bvh/triangle_bvh_tree.inl"
#include "scene/shape/triangle/triangle_primitive.hpp"

#include <chrono>
#include <cstdint>
#include <cstring>
#include <vector>

using scene::shape::Node_stack;
using scene::shape::Vertex_stream_separate_compact;
using scene::shape::triangle::Index_triangle;
using scene::shape::triangle::Intersection;
using scene::shape::triangle::bvh::Builder_SAH;
using scene::shape::triangle::bvh::Tree;

namespace {

struct Handle {
    Tree     tree;
    uint32_t num_nodes = 0;
    double   build_seconds = 0.0;
};

}  // namespace

extern "C" {

// Builds sprout's SBVH (16 slices, sweep threshold 64, 4 primitives per leaf,
// the settings of triangle_mesh_provider.cpp:377) over the given mesh.
void* sref_build(float const* positions, uint32_t num_vertices, uint32_t const* indices,
                 uint32_t num_triangles, uint32_t num_slices, uint32_t sweep_threshold,
                 uint32_t max_primitives, uint32_t num_threads) {
    std::vector<packed_float3> p(num_vertices);
    std::vector<packed_float3> n(num_vertices);
    for (uint32_t i = 0; i < num_vertices; ++i) {
        p[i] = packed_float3(positions[3 * i + 0], positions[3 * i + 1], positions[3 * i + 2]);
        n[i] = packed_float3(0.f, 0.f, 1.f);
    }

    Vertex_stream_separate_compact vertices(num_vertices, p.data(), n.data());

    std::vector<Index_triangle> triangles(num_triangles);
    for (uint32_t i = 0; i < num_triangles; ++i) {
        triangles[i] = Index_triangle(indices[3 * i + 0], indices[3 * i + 1], indices[3 * i + 2], i);
    }

    thread::Pool threads(num_threads);

    auto* h = new Handle;

    auto const start = std::chrono::steady_clock::now();

    Builder_SAH builder(num_slices, sweep_threshold, max_primitives);
    builder.build(h->tree, num_triangles, triangles.data(), vertices, threads);

    h->build_seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    h->num_nodes = h->tree.num_nodes_;

    return h;
}

void sref_free(void* handle) {
    delete static_cast<Handle*>(handle);
}

double sref_build_seconds(void* handle) {
    return static_cast<Handle*>(handle)->build_seconds;
}

uint32_t sref_num_nodes(void* handle) {
    return static_cast<Handle*>(handle)->num_nodes;
}

uint32_t sref_num_slots(void* handle) {
    return static_cast<Handle*>(handle)->tree.num_triangles();
}

// 32 bytes per node, the in-memory layout of scene::bvh::Node
// (scene_bvh_node.hpp:44-59).
void sref_nodes(void* handle, void* out) {
    auto* h = static_cast<Handle*>(handle);
    std::memcpy(out, h->tree.nodes_, size_t(h->num_nodes) * sizeof(scene::bvh::Node));
}

// Input triangle id held by each tree slot, in slot order.
void sref_slots(void* handle, uint32_t* out) {
    auto* h = static_cast<Handle*>(handle);
    for (uint32_t i = 0, len = h->tree.num_triangles(); i < len; ++i) {
        out[i] = h->tree.triangle_part(i);
    }
}

// rays: 8 floats each {o.x, o.y, o.z, min_t, d.x, d.y, d.z, max_t}.
// any == 0: hits gets 4 words per ray {t, u, v, input triangle id | 0xFFFFFFFF}.
// any != 0: hits gets 1 word per ray (1 occluded, 0 free).
// Returns the wall time of the tracing loop in seconds.
double sref_trace(void* handle, float const* rays, uint64_t num_rays, int32_t any,
                  uint32_t* hits, uint32_t num_threads) {
    auto* h = static_cast<Handle*>(handle);

    Tree const& tree = h->tree;

    thread::Pool threads(num_threads);

    auto const start = std::chrono::steady_clock::now();

    threads.run_range(
        [&tree, rays, any, hits](uint32_t /*id*/, int32_t begin, int32_t end) noexcept {
            Node_stack stack;

            for (int32_t r = begin; r < end; ++r) {
                float const* ray = rays + 8 * uint64_t(r);

                // per query, as BVH_wrapper does (prop_bvh_wrapper.cpp:40,186)
                stack.clear();

                Simdf const origin(_mm_setr_ps(ray[0], ray[1], ray[2], ray[3]));
                Simdf const direction(_mm_setr_ps(ray[4], ray[5], ray[6], ray[7]));

                scalar const min_t(ray[3]);
                scalar       max_t(ray[7]);

                if (any) {
                    hits[r] = tree.intersect_p(origin, direction, min_t, max_t, stack) ? 1u : 0u;
                    continue;
                }

                Intersection isec;

                uint32_t* out = hits + 4 * uint64_t(r);

                if (tree.intersect(origin, direction, min_t, max_t, stack, isec)) {
                    float const t = max_t.x();
                    float const u = isec.u.x();
                    float const v = isec.v.x();
                    std::memcpy(out + 0, &t, 4);
                    std::memcpy(out + 1, &u, 4);
                    std::memcpy(out + 2, &v, 4);
                    out[3] = tree.triangle_part(isec.index);
                } else {
                    float const t = max_t.x();
                    std::memcpy(out + 0, &t, 4);
                    out[1] = 0;
                    out[2] = 0;
                    out[3] = 0xFFFFFFFFu;
                }
            }
        },
        0, int32_t(num_rays));

    return std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
}

// The reference's math::transform_point / transform_vector (matrix4x4.inl:
// 319-342) and AABB::transform (aabb.inl:263-280) as the reference build
// compiles them.  m: 16 floats, row-vector rows r0..r3.
void sref_transform_point(float const* m, float const* v, float* out) {
    math::Simd4x4f const mm(*reinterpret_cast<float4x4 const*>(m));
    float3 const         r(math::transform_point(mm, Simdf(_mm_setr_ps(v[0], v[1], v[2], 0.f))));
    out[0] = r[0];
    out[1] = r[1];
    out[2] = r[2];
}

void sref_transform_vector(float const* m, float const* v, float* out) {
    math::Simd4x4f const mm(*reinterpret_cast<float4x4 const*>(m));
    float3 const         r(math::transform_vector(mm, Simdf(_mm_setr_ps(v[0], v[1], v[2], 0.f))));
    out[0] = r[0];
    out[1] = r[1];
    out[2] = r[2];
}

void sref_transform_box(float const* m, float const* lo, float const* hi, float* out) {
    AABB const b(float3(lo[0], lo[1], lo[2]), float3(hi[0], hi[1], hi[2]));
    AABB const t = b.transform(*reinterpret_cast<float4x4 const*>(m));
    for (int k = 0; k < 3; ++k) {
        out[k]     = t.bounds[0][k];
        out[3 + k] = t.bounds[1][k];
    }
}

// scene::bvh::Builder (scene_bvh_builder.cpp:14-107) over n prop boxes (6
// floats each: lo, hi), as Scene::compile builds the prop tree
// (scene.cpp:235).  nodes (32 B each) / indices may be null to query the
// counts.
int32_t sref_build_props(float const* boxes, uint32_t n, uint32_t num_threads, void* nodes,
                         uint32_t* indices, uint32_t* num_nodes, uint32_t* num_indices) {
    std::vector<uint32_t> ids(n);
    std::vector<AABB>     aabbs(n);
    for (uint32_t i = 0; i < n; ++i) {
        ids[i]   = i;
        aabbs[i] = AABB(float3(boxes[6 * i + 0], boxes[6 * i + 1], boxes[6 * i + 2]),
                        float3(boxes[6 * i + 3], boxes[6 * i + 4], boxes[6 * i + 5]));
    }
    thread::Pool        threads(num_threads ? num_threads : 1);
    scene::bvh::Builder builder;
    scene::bvh::Tree    tree;
    builder.build(tree, ids, aabbs, threads);
    *num_nodes   = tree.num_nodes_;
    *num_indices = tree.num_indices_;
    if (nodes) {
        std::memcpy(nodes, tree.nodes_, size_t(tree.num_nodes_) * 32);
    }
    if (indices) {
        std::memcpy(indices, tree.indices_, size_t(tree.num_indices_) * 4);
    }
    return 0;
}

}  // extern "C"
