/*
 * sprout_gpu: MI355X (gfx950) SBVH build and ray traversal for sprout.
 *
 * C ABI.  Conventions follow sprout's own C API (source/capi/sprout.h:39-122,
 * source/capi/capi.cpp:112-131): int32_t status (0 ok, negative error),
 * resource handles are uint32_t with 0xFFFFFFFF meaning "none", no
 * exceptions, one library context per process, calls serialized by the
 * caller.  Ray/hit buffers are DEVICE pointers owned by the caller; kernels
 * are enqueued on the caller's hipStream_t (passed as void*, NULL = the null
 * stream) and complete asynchronously unless stated otherwise.
 */
#ifndef SPROUT_GPU_H
#define SPROUT_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MI_NULL 0xFFFFFFFFu

enum mi_status {
    MI_OK              = 0,
    MI_ERR_NO_CONTEXT  = -1, /* as su_*: call before mi_init (capi.cpp:112-123) */
    MI_ERR_ARGS        = -2, /* invalid arguments */
    MI_ERR_HIP         = -3, /* HIP runtime error (device allocation, launch) */
    MI_ERR_HANDLE      = -4, /* unknown bvh handle */
    MI_ERR_UNSPLITTABLE = -5, /* a node of more than 255 references could not be split
                                 (the reference logs "Cannot split node further",
                                 scene_bvh_builder_base.cpp:50-57) */
    MI_ERR_STACK       = -6, /* a ray needed more than 127 traversal stack entries
                                 (node_stack.hpp:24) */
};

/* One ray, 32 bytes: the first two float4 of sprout's math::ray
 * (base/math/ray.hpp; ray.inl:11-32): origin.xyz + min_t, direction.xyz + max_t. */
typedef struct mi_ray {
    float origin[3];
    float min_t;
    float direction[3];
    float max_t;
} mi_ray;

/* Closest-hit result, 16 bytes.  t is the ray's max_t after the query (the
 * hit distance, or the unchanged max_t on a miss); u, v the barycentrics
 * (triangle_bvh_tree.hpp:32-37); triangle the INPUT triangle index, or
 * MI_NULL on a miss. */
typedef struct mi_hit {
    float    t;
    float    u;
    float    v;
    uint32_t triangle;
} mi_hit;

/* SAH builder settings; sprout meshes use {16, 64, 4}
 * (triangle_mesh_provider.cpp:377, Builder_SAH(num_slices, sweep_threshold,
 * max_primitives)). */
typedef struct mi_bvh_settings {
    uint32_t num_slices;
    uint32_t sweep_threshold;
    uint32_t max_primitives;
} mi_bvh_settings;

typedef struct mi_bvh_info {
    uint32_t num_nodes;
    uint32_t num_slots;     /* triangle references in leaves (>= triangles with spatial splits) */
    uint32_t num_triangles;
    uint32_t max_depth;     /* deepest node, root = 0 */
    float    aabb_min[3];
    float    aabb_max[3];
    double   build_ms;      /* device time of the build, 0 for an upload */
} mi_bvh_info;

/* Library context on one device.  Replaces su_init (capi.cpp:137) for the GPU path. */
int32_t mi_init(int32_t device);

/* Frees every bvh and the context.  su_release (capi.cpp:150). */
int32_t mi_release(void);

/* Text of the last error of this thread ("" if none). */
char const* mi_last_error(void);

/*
 * Builds sprout's SBVH on the device from a device-resident indexed mesh:
 * positions = num_vertices x 3 floats with a stride of position_stride floats
 * (>= 3), indices = num_triangles x 3 uint32.  The tree is the one
 * scene::shape::triangle::bvh::Builder_SAH::build produces
 * (triangle_bvh_builder_sah.cpp:21-67, scene_bvh_builder_base.cpp:26-290):
 * same nodes, same depth-first layout, same leaf order.  Synchronous.
 */
int32_t mi_bvh_build(float const* positions, uint32_t num_vertices, uint32_t position_stride,
                     uint32_t const* indices, uint32_t num_triangles,
                     mi_bvh_settings const* settings, void* stream, uint32_t* bvh);

/*
 * Uploads a tree already built on the host in sprout's serialized layout:
 * nodes = num_nodes x 32 bytes (scene::bvh::Node, scene_bvh_node.hpp:44-59),
 * slots = num_slots input-triangle indices in tree-slot order (what
 * Tree::allocate_triangles/set_triangle fill, triangle_bvh_tree.inl:87-96).
 * All pointers are HOST pointers.  Synchronous.
 */
int32_t mi_bvh_upload(void const* nodes, uint32_t num_nodes, uint32_t const* slots,
                      uint32_t num_slots, float const* positions, uint32_t num_vertices,
                      uint32_t position_stride, uint32_t const* indices, uint32_t num_triangles,
                      uint32_t* bvh);

int32_t mi_bvh_release(uint32_t bvh);

int32_t mi_bvh_get_info(uint32_t bvh, mi_bvh_info* info);

/* Copies the tree back to HOST memory in the same layout as mi_bvh_upload
 * takes (nodes: num_nodes x 32 bytes, slots: num_slots words).  Synchronous. */
int32_t mi_bvh_download(uint32_t bvh, void* nodes, uint32_t* slots);

/*
 * Closest hit for num_rays device-resident rays.
 * Replaces bvh::Tree::intersect (triangle_bvh_tree.cpp:41-90): identical
 * visiting order, arithmetic and tie rule, with the reciprocal pinned to
 * rcp(x) = correctly rounded 1/x followed by the reference's Newton step.
 */
int32_t mi_trace_closest(uint32_t bvh, mi_ray const* rays, uint32_t num_rays, mi_hit* hits,
                         void* stream);

/* Any hit: occluded[i] = 1 if ray i hits anything in [min_t, max_t], else 0.
 * Replaces bvh::Tree::intersect_p (triangle_bvh_tree.cpp:145-185). */
int32_t mi_trace_any(uint32_t bvh, mi_ray const* rays, uint32_t num_rays, uint32_t* occluded,
                     void* stream);

/*
 * Work counters: for each ray, counts[2i] = nodes the reference's loop fetches
 * and counts[2i+1] = triangle tests it performs (any_hit selects the
 * intersect_p loop).  Also returns the closest hits / occlusion in hits_or_occluded
 * (mi_hit* or uint32_t*).  Used to derive algorithmic bytes per ray.
 */
int32_t mi_trace_counts(uint32_t bvh, mi_ray const* rays, uint32_t num_rays, int32_t any_hit,
                        void* hits_or_occluded, uint32_t* counts, void* stream);

/* Traversal kernel launch configuration (0 = library default). */
typedef struct mi_trace_config {
    uint32_t block_size;       /* threads per workgroup, multiple of 64 (default 256) */
    uint32_t blocks_per_cu;    /* workgroups per CU of the persistent grid (default 8) */
    uint32_t kernel;           /* 0: per-lane ray refill, 1: static grid-stride,
                                  2: refill with inner/leaf phases (while-while, default) */
    uint32_t refill_lanes;     /* kernels 0, 2: hand out rays once this many lanes are idle (24) */
    uint32_t claim_batch;      /* kernels 0, 2: rays a wave claims per atomic (64) */
    uint32_t lds_entries;      /* stack entries per lane kept in LDS, <= 32 (default 16;
                                  0: up to 32); deeper entries spill to global memory */
} mi_trace_config;

int32_t mi_set_trace_config(mi_trace_config const* config);

/* Waits for the stream and reports errors of earlier calls on it.
 * Traversals share the context's ray counter and stack spill area: calls on
 * different streams must not run at the same time. */
int32_t mi_sync(void* stream);

/* ---- synthetic scenes (host-side, deterministic; used by tests and bench) ---- */

/* Icosphere of num_subdivisions subdivisions of an icosahedron (20 * 4^s
 * triangles, 10 * 4^s + 2 vertices), radius 1, radially displaced by
 * amplitude * fBm (octaves of value noise), then num_random_triangles small
 * triangles in the unit cube with log-normal(mu, sigma) edge lengths.
 * Pass NULL buffers to query the counts.  HOST pointers. */
int32_t mi_gen_displaced_icosphere(uint32_t num_subdivisions, uint32_t octaves, float amplitude,
                                   uint64_t noise_seed, uint32_t num_random_triangles,
                                   float log_mu, float log_sigma, uint64_t random_seed,
                                   float* positions, uint32_t* num_vertices, uint32_t* indices,
                                   uint32_t* num_triangles);

/* Pinhole camera rays, width x height, one per pixel centre, row-major;
 * vertical field of view in degrees.  HOST pointer, width*height rays. */
int32_t mi_gen_camera_rays(uint32_t width, uint32_t height, float const eye[3],
                           float const target[3], float const up[3], float fov_deg,
                           mi_ray* rays);

/* Cosine-distributed bounce rays from closest hits: ray i starts at the hit
 * point of the (i mod m)-th hitting input ray (m = number of hits), offset
 * along the geometric normal facing the incoming ray with offset_ray
 * (core/scene/ray_offset.inl:18-29), direction cosine-distributed about that
 * normal.  HOST pointers.  Returns MI_ERR_ARGS if no input ray hits. */
int32_t mi_gen_bounce_rays(mi_ray const* in_rays, mi_hit const* hits, uint32_t num_in,
                           float const* positions, uint32_t position_stride,
                           uint32_t const* indices, uint64_t seed, mi_ray* out_rays,
                           uint32_t num_out);

#ifdef __cplusplus
}
#endif

#endif
