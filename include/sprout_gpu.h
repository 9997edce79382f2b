/*
 * sprout_gpu: MI355X (gfx950) SBVH build, ray traversal and wavefront MIS
 * path tracing for sprout.
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

/* Closest hit of an instanced scene: t, barycentrics u / v and the
 * triangle (input index in its mesh) as mi_hit, plus the instance. */
typedef struct mi_instance_hit {
    float    t;
    float    u;
    float    v;
    uint32_t triangle;
    uint32_t instance; /* MI_NULL on a miss */
    uint32_t slot;     /* the triangle's slot in its mesh tree (Intersection::index) */
    uint32_t pad[2];
} mi_instance_hit;

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

/*
 * mi_trace_closest without device allocation.  No counterpart in sprout: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 * The ray counter and the traversal stack's spill area come from a fixed
 * 32 MiB device array of the library (each GPU holds its own until the
 * process ends; mi_init looks up its address).  The results equal
 * mi_trace_closest's; the persistent grid shrinks to what the array holds.
 * MI_ERR_ARGS when the tree's stack (max_depth + 2 entries, those beyond
 * the LDS part spilled) needs more than the array for one workgroup.
 * Calls that overlap on one GPU (other streams or threads) race on the array.
 */
int32_t mi_trace_closest_noalloc(uint32_t bvh, mi_ray const* rays, uint32_t num_rays, mi_hit* hits,
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
                                  2: refill with inner/leaf phases (while-while),
                                  3: as 2 with stack entries that carry the child's node word
                                  and box entry distance (default; trees of < 2^22 nodes and
                                  slots, else 2),
                                  4: as 3 over a compressed copy of the tree: 32 bytes per
                                  sibling pair (16-bit grid boxes that hold the originals, the
                                  node words), half the bytes of the 64-byte pair; the same
                                  nodes visited in the same order, plus any the grown boxes
                                  admit, so the same hits */
    uint32_t refill_lanes;     /* kernels 0, 2, 3: hand out rays once this many lanes are idle (8) */
    uint32_t claim_batch;      /* kernels 0, 2, 3: rays a wave claims per atomic (64) */
    uint32_t lds_entries;      /* stack entries per lane kept in LDS, <= 32 (default 10;
                                  0: up to 32); deeper entries spill to global memory */
    uint32_t sort_rays;        /* kernels 3 and 4, mi_trace_closest / mi_trace_any: trace the
                                  rays in the order of (direction octant, Morton code of the
                                  origin), a radix sort per call; results in the callers'
                                  order, the same values (0: off) */
} mi_trace_config;

int32_t mi_set_trace_config(mi_trace_config const* config);

/* Autotuner of the traversal's launch shape: times closest-hit traversals of
 * the given DEVICE rays through the tree over candidate shapes (kernels 3
 * and 4; LDS stack entries, refill threshold, claim batch and workgroups per
 * CU around the defaults, then the best with the rays sorted), `repeats` timed launches each after one warm-up,
 * with HIP events on `stream`; makes the fastest the context's shape (as
 * mi_set_trace_config) and returns it in *best (HOST, may be NULL) and its
 * mean launch time in *best_ms.  Every shape gives the same results, so
 * tuning changes only the speed.  Waits for the stream.  No counterpart in
 * sprout (its traversal has no launch shape). */
int32_t mi_trace_autotune(uint32_t bvh, mi_ray const* rays, uint32_t num_rays, uint32_t repeats,
                          void* stream, mi_trace_config* best, float* best_ms);

/* Waits for the stream and reports errors of earlier calls on it.
 * Traversals share the context's ray counter and stack spill area: calls on
 * different streams must not run at the same time. */
int32_t mi_sync(void* stream);

/* ---- wavefront MIS path tracer ---- */

enum mi_material_kind {
    MI_MATERIAL_SUBSTITUTE = 0, /* Disney diffuse + GGX with energy compensation
                                   (substitute_material.cpp:19-30, substitute_sample.cpp:10-45) */
    MI_MATERIAL_LAMBERT    = 1, /* lambert::Iso (lambert/lambert.inl:15-45) */
    MI_MATERIAL_LIGHT      = 2, /* one-sided constant emitter, pure emissive
                                   (light/light_constant.cpp:13-35) */
};

/* A material.  emission is the radiance of a LIGHT material; ior is the
 * Substitute's index of refraction (f0 = schlick_f0(ior, 1)). */
typedef struct mi_material {
    float    color[3];
    float    emission[3];
    float    roughness;
    float    metallic;
    float    ior;
    uint32_t kind;
} mi_material;

/* sprout's pinhole Perspective camera (camera_perspective.cpp:47-80,
 * 204-226): the camera entity's world position and the rows of its rotation
 * (r0 the image's right, r1 its up, r2 the view direction; sprout's
 * Composed_transformation rotation with unit scale) and the HORIZONTAL field
 * of view in degrees (the take's "fov").  Pixel (x, y), row-major from the
 * top-left, with pixel sample (u, v) looks along
 *   normalize(left_top + (x + u) d_x + (y + v) d_y) . [r0; r1; r2],
 * left_top = (-1, h / w, 1 / tan(fov / 2)), d_x = (2 / w, 0, 0),
 * d_y = (0, -2 (h / w) / h, 0), computed in fp32 as sprout's on_update does. */
typedef struct mi_camera {
    uint32_t width;
    uint32_t height;
    float    position[3];
    float    rotation[9]; /* rows r0, r1, r2 */
    float    fov_deg;     /* horizontal */
} mi_camera;

/* Pathtracer_MIS settings (pathtracer_mis.hpp Settings; take_loader.cpp:617-739).
 * A render runs samples_per_pixel samples as waves of samples_per_wave
 * samples of every pixel (samples_per_pixel must be a multiple of it).
 * Sample s of pixel (x, y) draws its random numbers from sprout's PCG32
 * (base/random/generator.inl:16-49) started as start(seed, s * W * H +
 * y * W + x), like the per-pixel start of rendering_worker.cpp:104 with the
 * sample index in place of the iteration; samples first_sample ..
 * first_sample + samples_per_pixel - 1 are rendered, so successive calls
 * continue a progressive render. */
typedef struct mi_render_settings {
    uint32_t samples_per_pixel;
    uint32_t samples_per_wave;
    uint32_t first_sample;
    uint32_t min_bounces; /* Russian roulette from this depth on (sprout default 4) */
    uint32_t max_bounces; /* <= 64 */
    uint32_t flags;       /* MI_RENDER_* */
    uint64_t seed;
} mi_render_settings;

enum mi_render_flags {
    MI_RENDER_NO_GRAPH       = 1, /* launch the bounce loop kernel by kernel instead of as a hipGraph */
    MI_RENDER_NO_LIGHT_SPLIT = 2, /* one light per vertex (sprout's Light_sampling::Single); by
                                     default the light tree splits at depths 0-2 (Adaptive,
                                     take_loader.cpp:624, pathtracer_mis.cpp:423-426) */
    MI_RENDER_GOLDEN_RATIO   = 4, /* pixel samples from sprout's Golden_ratio sampler
                                     (sampler_golden_ratio.cpp:21-73): per pixel, the generator
                                     started as Worker::render starts it (seed 0, sequence
                                     (y + r) W' + x + r over the grid grown by the filter reach
                                     r, rendering_worker.cpp:97-104), a golden-ratio sequence
                                     of samples_per_pixel points biased-shuffled, sample s
                                     takes point s; otherwise two draws of the path's own
                                     generator.  The frame is the one call: first_sample must
                                     be 0 */
    MI_RENDER_REGENERATE     = 8, /* regenerating waves (mi_render with the box sensor): a path
                                     slot whose path ends starts the frame's next sample, so the
                                     ray queues stay full while paths of uneven depth finish; the
                                     image equals the wave-by-wave render bit for bit (each
                                     sample's random numbers depend on the sample alone, and the
                                     samples are summed per pixel in the same order).  Needs room
                                     for min(spp, 2^27 / pixels) planes of samples (12 bytes each) */
    MI_RENDER_SPECTRAL       = 16, /* hero-wavelength transport (no counterpart in sprout, which
                                     renders RGB): every path carries four wavelengths over
                                     380-720 nm, a hero drawn by a hash of (seed, the path's
                                     generator sequence) and three at quarter-range offsets;
                                     rgb colours, emission and medium coefficients are
                                     upsampled by a partition of unity of three bands (blue
                                     falling 450-510 nm, red rising 560-610 nm, green the
                                     rest), tracking and transmittance run per wavelength with
                                     history-aware weights, and the film reads rgb through the
                                     bands' dual basis.  Grey scenes render the rgb image in
                                     expectation, with the same random numbers per path */
};

typedef struct mi_render_stats {
    double seconds;      /* host wall time of the call, launch to completion */
    double paths;        /* camera samples */
    double extend_rays;  /* closest-hit rays traced */
    double shadow_rays;  /* any-hit rays traced */
    double trained;      /* guiding samples trained on (mi_render_guided) */
} mi_render_stats;

/* Creates a scene from HOST arrays: triangles with one material each; every
 * emissive triangle (kind LIGHT) is one light of the scene's light tree
 * (light::Tree, core/scene/light/light_tree.cpp, built on the host as
 * Tree_builder does, light_tree_builder.cpp:436-540).  The tree is built on the device with *bvh_settings (NULL:
 * {16, 64, 4}).  Mirrors su_create_triangle_mesh_async + material creation +
 * Scene::compile (capi.cpp:378-406, scene.cpp:216-267). */
int32_t mi_scene_create(float const* positions, uint32_t num_vertices, uint32_t const* indices,
                        uint32_t num_triangles, uint32_t const* triangle_materials,
                        mi_material const* materials, uint32_t num_materials,
                        mi_bvh_settings const* bvh_settings, uint32_t* scene);

/* A mesh for mi_scene_create_instanced (HOST arrays). */
typedef struct mi_mesh {
    float const*    positions; /* 3 floats per vertex */
    uint32_t        num_vertices;
    uint32_t const* indices;   /* 3 per triangle */
    uint32_t        num_triangles;
} mi_mesh;

/* Creates an instanced scene: num_instances placements of the meshes
 * (instance i shows meshes[instance_mesh[i]] with the row-vector 4x4
 * matrices world_to_object / object_to_world, as mi_instances_create, and
 * the one material instance_materials[i], which may not be emissive), plus
 * the world-space triangles of positions / indices with their materials as in
 * mi_scene_create (the emissive ones form the light tree; at least one
 * triangle).  Each mesh gets its own tree, and the top-level tree spans the
 * instances and the world triangles (one more instance, identity matrices).
 * The hit point of an instance is interpolated in object space and
 * transformed to world space, its normal transformed and normalized
 * (Mesh::intersect, triangle_mesh.cpp:364-392; similarity transforms).
 * Mirrors su_create_triangle_mesh_async + su_create_prop with one material +
 * su_entity_set_transformation (capi.cpp:378-456, 521-541) and the prop tree of
 * Scene::compile (scene.cpp:216-267).  HOST arrays throughout. */
int32_t mi_scene_create_instanced(mi_mesh const* meshes, uint32_t num_meshes,
                                  uint32_t const* instance_mesh, float const* world_to_object,
                                  float const* object_to_world,
                                  uint32_t const* instance_materials, uint32_t num_instances,
                                  float const* positions, uint32_t num_vertices,
                                  uint32_t const* indices, uint32_t num_triangles,
                                  uint32_t const* triangle_materials, mi_material const* materials,
                                  uint32_t num_materials, mi_bvh_settings const* bvh_settings,
                                  uint32_t* scene);

/* Moves the instances of an instanced scene: new world_to_object /
 * object_to_world matrices (HOST, 16 floats each, one pair per instance in
 * creation order, num_instances as at creation).  Rebuilds the top-level tree
 * on the device from the instance boxes and updates the transforms the hit
 * points use; meshes, their trees, materials and the light tree (its
 * emitters are the world triangles) are kept.  A render afterwards equals
 * one of a scene created with the new matrices.  Waits for work in flight;
 * on failure the scene keeps its old placement.  Mirrors
 * su_entity_set_transformation (capi.cpp:521-541) followed by the per-frame
 * prop-tree rebuild of Scene::compile (scene.cpp:216-267). */
int32_t mi_scene_update_instances(uint32_t scene, float const* world_to_object,
                                  float const* object_to_world, uint32_t num_instances);

/* New vertex positions for a scene of mi_scene_create (same vertex count and
 * topology; positions: DEVICE array, 3 floats per vertex), as a frame of an
 * animation moves its props and emitters: everything rebuilds on the device
 * on `stream` (the call waits): each emissive triangle's light, every
 * emissive part's Primitive_tree and cone, the scene's light tree over the
 * parts (Scene::compile, scene.cpp:216-267, light_tree_builder.cpp:436-760),
 * the same trees bit for bit that mi_scene_create would build for these
 * positions, then the tree over the triangles with the GPU builder.  A scene
 * light-tree node that only the host builder's median split divides is built
 * on the host.  MI_ERR_ARGS, and the old positions kept, if an emissive
 * triangle would lose its area.  light_ms (optional): host wall time of the
 * light part.  Waits for renders in flight.  Instanced scenes move their
 * instances with mi_scene_update_instances.  Both updates keep the scene
 * box of its creation as the guide's position normalization, so a trained
 * guide sees the same inputs after a move (points beyond it clamp). */
int32_t mi_scene_update_positions(uint32_t scene, float const* positions, uint32_t num_vertices,
                                  double* light_ms, void* stream);

int32_t mi_scene_release(uint32_t scene);

/* Renders and ADDS each pixel's radiance sum over the samples into accum,
 * a DEVICE buffer of width * height * 3 doubles (rgb); the image is accum
 * divided by the total sample count.  Enqueued on stream (void*
 * hipStream_t, NULL = null stream).  stats (HOST, may be NULL) makes the
 * call wait for the render and fill it.  Replaces the per-sample
 * Pathtracer_MIS::li (pathtracer_mis.cpp:59-93) behind the tile loop of
 * Worker::render (rendering_worker.cpp:73-127) for a whole frame, the
 * boundary of Driver::render_frame_forward (rendering_driver.cpp:295-360). */
int32_t mi_render(uint32_t scene, mi_camera const* camera, mi_render_settings const* settings,
                  double* accum, void* stream, mi_render_stats* stats);

/* ---- filtered sensor ---- */

enum mi_filter_kind {
    MI_FILTER_BOX      = 0, /* the unfiltered sensor: a sample counts for its own pixel */
    MI_FILTER_GAUSSIAN = 1, /* filter::Gaussian(radius, alpha) */
    MI_FILTER_BLACKMAN = 2, /* filter::Blackman(radius) */
    MI_FILTER_MITCHELL = 3  /* filter::Mitchell(radius, b, c) */
};

/* sprout's sensor (take "sensor": {"filter": ..., "clamp": ...}, take_loader.cpp:
 * 331-436): filter radius in (0, 2] (Filtered_1p0 up to 1, Filtered_2p0 up to 2;
 * take defaults radius 2, alpha 1.8, b = c = 1/3).  clamp_max[0] < 0: no clamp;
 * clamp_max[1] < 0: clamp::Luminance(clamp_max[0]); else clamp::Clamp(clamp_max). */
typedef struct mi_sensor {
    uint32_t filter; /* mi_filter_kind */
    float    radius;
    float    alpha;
    float    b;
    float    c;
    float    clamp_max[3];
} mi_sensor;

/* As mi_render through sprout's filtered sensor: the camera samples pixels of
 * the image grown by the filter's reach on every side (the worker's extended
 * tiles, rendering_worker.cpp:93-104; the generator of grid pixel (gx, gy) is
 * started at sample * grid pixels + gy * grid width + gx) and every sample
 * adds weight * clamped colour and weight to the image pixels within reach
 * (Filtered_1p0 / Filtered_2p0::add_sample, filtered.inl:121-300).  ADDS to
 * accum, a DEVICE buffer of width * height * 4 doubles (weighted rgb,
 * weight); the image is rgb / weight (Opaque::resolve, opaque.cpp:61-69). */
int32_t mi_render_filtered(uint32_t scene, mi_camera const* camera,
                           mi_render_settings const* settings, mi_sensor const* sensor,
                           double* accum, void* stream, mi_render_stats* stats);

/* ---- instanced meshes (sprout's props) ---- */

/* A top-level tree over num_instances instances of mesh trees: instance i
 * places the tree meshes[i] (a handle of mi_bvh_build / mi_bvh_upload; the
 * same mesh may be placed many times) with the row-vector 4x4 matrices of
 * sprout's Composed_transformation, world_to_object and object_to_world (16
 * floats each, rows r0..r3, p' = x r0 + y r1 + z r2 + r3; HOST).  Its world
 * box is the mesh's box transformed as AABB::transform does (aabb.inl:263-280);
 * the top-level tree is the SBVH scene::bvh::Builder builds over those boxes
 * (scene_bvh_builder.cpp:14-107, settings as mi_bvh_build's).  The mesh
 * trees must outlive it.  Mirrors su_create_prop + su_entity_set_transformation
 * (capi/capi.cpp:521-541) and the prop tree of Scene::compile (scene.cpp:235). */
int32_t mi_instances_create(uint32_t const* meshes, float const* world_to_object,
                            float const* object_to_world, uint32_t num_instances,
                            mi_bvh_settings const* settings, void* stream, uint32_t* instances);
int32_t mi_instances_release(uint32_t instances);

/* Top-level then mesh traversal of DEVICE rays (world space), as
 * BVH_wrapper::intersect / intersect_p (prop_bvh_wrapper.cpp:33-106,
 * 183-247) over Prop::intersect (prop.cpp:68-126, the prop box first) and
 * Mesh::intersect (triangle_mesh.cpp:366-446, the ray in object space,
 * same t).  hits: n mi_instance_hit; occluded: n uint32 (1 = hit). */
int32_t mi_trace_instances_closest(uint32_t instances, mi_ray const* rays, uint32_t num_rays,
                                   mi_instance_hit* hits, void* stream);
int32_t mi_trace_instances_any(uint32_t instances, mi_ray const* rays, uint32_t num_rays,
                               uint32_t* occluded, void* stream);
/* HOST: the top-level tree (nodes: 32-byte sprout nodes, slots: instance per
 * slot) and each instance's world box (6 floats: lo, hi).  NULL arrays query
 * *num_nodes / *num_slots. */
int32_t mi_instances_download(uint32_t instances, void* nodes, uint32_t* slots, float* boxes,
                              uint32_t* num_nodes, uint32_t* num_slots);

/* ---- light tree ---- */

/* One node of the light tree, 48 bytes: light::Node (light_tree.hpp:32-65). */
typedef struct mi_light_node {
    float    center[3];
    float    radius;
    float    cone[4]; /* axis, cos of the opening angle */
    float    power;
    float    variance;
    uint32_t bits; /* has_children : 1, two_sided : 1, children_or_light : 30 */
    uint32_t num_lights;
} mi_light_node;

/* Builds the light tree of num_lights lights on the host, as Tree_builder
 * builds the scene tree (light_tree_builder.cpp:436-540, finite lights):
 * per light a box (min, max: 3 floats each), a cone (4 floats), a power and
 * a two-sided flag.  HOST arrays; nodes needs 2 * num_lights - 1 entries
 * (NULL: query *num_nodes only); middles as many; mapping (order -> light)
 * and orders (light -> order) num_lights each.  No GPU needed. */
int32_t mi_light_tree_build(float const* box_min, float const* box_max, float const* cones,
                            float const* powers, uint8_t const* two_sided, uint32_t num_lights,
                            mi_light_node* nodes, uint32_t* middles, uint32_t* mapping,
                            uint32_t* orders, uint32_t* num_nodes);

/* Builds the Primitive_tree of one emissive part on the host, as
 * Tree_builder builds it (light_tree_builder.cpp:539-795): its num_lights
 * triangles as lights (box, cone, power = area; one-sided), the part's cone
 * (4 floats: dominant axis, cos of the widest angle), leaves of up to
 * max(num_lights / 64, 4) triangles.  Arrays as mi_light_tree_build; cdf_offset
 * (one per node: the offset of the leaf's distribution in cdfs, or 0xFFFFFFFF)
 * and cdfs (*num_cdfs floats, NULL: query) for leaves of more than 4 lights. */
int32_t mi_light_part_tree_build(float const* box_min, float const* box_max, float const* cones,
                                 float const* powers, uint32_t num_lights, float const* part_cone,
                                 mi_light_node* nodes, uint32_t* middles, uint32_t* mapping,
                                 uint32_t* orders, uint32_t* cdf_offset, float* cdfs,
                                 uint32_t* num_nodes, uint32_t* num_cdfs);

/* As mi_light_tree_build, on the device: DEVICE arrays, the same tree bit for
 * bit (nodes, middles, mapping, orders), built a level of nodes at a time on
 * `stream` (the call waits for it).  *num_nodes (host) receives the node
 * count.  MI_ERR_UNSPLITTABLE: a node every candidate leaves one side of,
 * which the host builder splits at the median; build on the host then.  No
 * counterpart in the reference, whose Tree_builder runs on its thread pool
 * (light_tree_builder.cpp:436-540; Scene::compile, scene.cpp:216-267). */
int32_t mi_light_tree_build_device(float const* box_min, float const* box_max, float const* cones,
                                   float const* powers, uint8_t const* two_sided,
                                   uint32_t num_lights, mi_light_node* nodes, uint32_t* middles,
                                   uint32_t* mapping, uint32_t* orders, uint32_t* num_nodes,
                                   void* stream);

/* The Primitive_trees of num_parts parts at once on the device, each as
 * mi_light_part_tree_build builds it: part r holds lights first[r] ..
 * first[r + 1] - 1 (first: HOST array of num_parts + 1, from 0 to
 * num_lights, no part empty) and its cone part_cones[4 r ..].  DEVICE
 * arrays; part r's nodes, middles and cdf_offset start at 2 first[r] - r,
 * its cdfs at 2 first[r]; mapping and orders are local to the part.
 * num_nodes: HOST, per part; num_cdfs: DEVICE, per part. */
int32_t mi_light_part_trees_build_device(float const* box_min, float const* box_max,
                                         float const* cones, float const* powers,
                                         uint32_t num_lights, uint32_t const* first,
                                         uint32_t num_parts, float const* part_cones,
                                         mi_light_node* nodes, uint32_t* middles,
                                         uint32_t* mapping, uint32_t* orders,
                                         uint32_t* cdf_offset, float* cdfs, uint32_t* num_nodes,
                                         uint32_t* num_cdfs, void* stream);

enum mi_light_flags {
    MI_LIGHT_SPLIT        = 1, /* split the tree (Adaptive light sampling) */
    MI_LIGHT_TOTAL_SPHERE = 2, /* the point is in a medium: no horizon test (translucent) */
};

/* Device queries of a scene's light tree (every emissive triangle is a light;
 * light ids number them in triangle order).  DEVICE arrays: points and
 * normals n x 3; split: MI_LIGHT_* flags.
 * sample: randoms n -> picks n x Max (64) x {light, pdf as float bits} and
 *         counts n (Tree::random_light, light_tree.cpp:363-450)
 * pdf:    lights n -> pdf n (Tree::pdf, light_tree.cpp:564-635)            */
int32_t mi_light_tree_sample(uint32_t scene, float const* points, float const* normals,
                             float const* randoms, uint32_t split, uint32_t n, uint32_t* picks,
                             uint32_t* counts, void* stream);
int32_t mi_light_tree_pdf(uint32_t scene, float const* points, float const* normals,
                          uint32_t const* lights, uint32_t split, uint32_t n, float* pdfs,
                          void* stream);
/* HOST: the scene's tree and lights, for checking (sizes from *num_nodes,
 * *num_lights; NULL arrays query them).  infos: per light center (3),
 * radius, cone (4), power, two_sided, triangle, area (12 words). */
int32_t mi_scene_light_tree(uint32_t scene, mi_light_node* nodes, uint32_t* middles,
                            uint32_t* mapping, uint32_t* orders, void* infos,
                            uint32_t* num_nodes, uint32_t* num_lights);

/* Shading normals of a scene's triangles: 3 floats per vertex (HOST, any
 * non-zero length; stored normalized).  A hit shades with its triangle's
 * vertex normals interpolated at the hit's barycentrics and normalized, as
 * sprout's Mesh::intersect does (triangle_mesh.cpp:398-414,
 * triangle_primitive_mt.inl:274-287); the geometric normal still decides
 * the side, offsets rays and weighs lights.  sprout gives every mesh vertex
 * a normal (the file's, or for a JSON mesh without them the last geometric
 * normal written to it, triangle_mesh_provider.cpp:103-118); without this
 * call a scene shades with the geometric normal.  normals NULL goes back to
 * that.  Moving the vertices (mi_scene_update_positions) keeps the normals.
 * In an instanced scene it covers the world triangles; meshes take theirs
 * through mi_scene_set_mesh_normals. */
int32_t mi_scene_set_normals(uint32_t scene, float const* normals, uint32_t num_vertices);

/* The same for mesh `mesh` of an instanced scene (mi_scene_create_instanced):
 * its vertex normals in object space (HOST, 3 per vertex), interpolated at a
 * hit in object space and moved to world space by the instance's
 * object-to-world matrix, as the geometric normal is (sprout turns both by
 * the prop's rotation, triangle_mesh.cpp:388-421: the same under uniform
 * scale).  normals NULL goes back to the geometric normal. */
int32_t mi_scene_set_mesh_normals(uint32_t scene, uint32_t mesh, float const* normals,
                                  uint32_t num_vertices);

/* ---- participating medium ---- */

/* Gives a scene a heterogeneous medium: a density grid of dims[0] x dims[1] x
 * dims[2] floats (HOST, x fastest, >= 0) stretched over the axis-aligned box
 * [box_lo, box_hi]; absorption cc_a and scattering cc_s per unit density
 * (rgb) and Henyey-Greenstein anisotropy g in (-1, 1).  Every ray segment in
 * the box is tracked (Tracking_multi, tracking_multi.cpp:44-183: delta
 * tracking with history-aware channel weights, majorants from the gridtree
 * Octree_builder builds, volumetric_octree_builder.cpp:13-296); shadow rays
 * through the box are attenuated by (residual) ratio tracking
 * (tracking.cpp:32-129).  Scattering vertices sample the phase function and
 * the light tree over the whole sphere.  Replaces the scene's medium;
 * density NULL removes it.  Mirrors the volumetric Grid material on a
 * box-shaped prop (volumetric_grid.cpp:18-70). */
int32_t mi_scene_set_medium(uint32_t scene, float const* density, uint32_t const dims[3],
                            float const box_lo[3], float const box_hi[3], float const cc_a[3],
                            float const cc_s[3], float g);

/* The majorants tracking steps with through a scene's medium. */
enum mi_majorants {
    MI_MAJORANTS_GRIDTREE = 0, /* sprout's gridtree (Octree_builder, volumetric_octree_builder.cpp:
                                  13-296): 32^3-voxel cells split into octants down to 8^3 voxels
                                  while their density range is 0.1 or more */
    MI_MAJORANTS_FLAT     = 1, /* a flat grid of 8^3-voxel cells (32^3 cells for 256^3 voxels),
                                  each with the minorant / majorant of its voxels and one voxel of
                                  border (the linear filter's reach); stored as a gridtree whose
                                  cells all split to their leaves.  No counterpart in sprout */
};

/* mi_scene_set_medium with a choice of majorants (MI_MAJORANTS_*); both give
 * unbiased tracking, they differ in how many null collisions a step makes. */
int32_t mi_scene_set_medium_majorants(uint32_t scene, float const* density, uint32_t const dims[3],
                                      float const box_lo[3], float const box_hi[3],
                                      float const cc_a[3], float const cc_s[3], float g,
                                      uint32_t majorants);

/* HOST, no GPU: the majorant gridtree of a density grid (Octree_builder::build).
 * num_cells (3): top-level cells of 32^3 voxels per axis.  nodes: uint32 per
 * node, bit 0 = has children, bits 1-31 = first child (8 consecutive) or
 * data index (0x7FFFFFFF: empty); data: 4 floats per leaf {minorant mu_a,
 * minorant mu_s, majorant mu_a, majorant mu_s}.  NULL arrays query the counts. */
int32_t mi_medium_gridtree(float const* density, uint32_t const dims[3], float const cc_a[3],
                           float const cc_s[3], uint32_t* num_cells, uint32_t* nodes,
                           uint32_t* num_nodes, float* data, uint32_t* num_data);

/* HOST, no GPU: as mi_medium_gridtree for either kind of majorants. */
int32_t mi_medium_majorant_tree(float const* density, uint32_t const dims[3], float const cc_a[3],
                                float const cc_s[3], uint32_t majorants, uint32_t* num_cells,
                                uint32_t* nodes, uint32_t* num_nodes, float* data,
                                uint32_t* num_data);

enum mi_medium_flags {
    MI_MEDIUM_TRANSMITTANCE = 1, /* Tracking::transmittance instead of Tracking_multi::integrate */
    MI_MEDIUM_NEAREST       = 2, /* unfiltered density lookups (Filter::Nearest) */
};

/* Batch tracking through a scene's medium, for checking: n DEVICE rays (each
 * clipped to the medium box as the renderer clips it), one PCG32 generator
 * per ray started at (seed, index), similarity-relation depth `depth`.
 * out: n x 5 DEVICE floats.
 * default: Tracking_multi::integrate -> {event (0 passed, 1 scattered,
 *   2 aborted), scattering t, weight rgb}
 * MI_MEDIUM_TRANSMITTANCE: Tracking::transmittance -> {1 transmitted /
 *   0 blocked, 0, tr rgb} */
int32_t mi_medium_track(uint32_t scene, mi_ray const* rays, uint32_t depth, uint64_t seed,
                        uint32_t flags, uint32_t n, float* out, void* stream);

/* The 32 x 32 GGX energy table E(n.wo, alpha) (rows alpha) regenerated on the
 * device by sprout's procedure (ggx_integrate.cpp:17-45, 202-240; the data of
 * ggx_integral.inl).  DEVICE pointer of 1024 floats. */
int32_t mi_ggx_e_table(float* table, void* stream);

/* Batch Substitute material queries on DEVICE arrays, one case per index:
 * params[6] (color rgb, roughness, metallic, ior), frames[12] (tangent,
 * bitangent, shading normal, geometric normal), wo[3].
 * evaluate: wi[3] -> out[4] (reflection * n.wi rgb, pdf)   (substitute_sample.cpp:10-26)
 * sample:   xi[3] (lobe choice, 2-D sample) -> out[8] (wi xyz, reflection rgb,
 *           pdf, bxdf type bits as uint32; pdf 0 = no sample) (substitute_sample.cpp:28-45) */
int32_t mi_substitute_evaluate(uint32_t n, float const* params, float const* frames,
                               float const* wo, float const* wi, float* out, void* stream);
int32_t mi_substitute_sample(uint32_t n, float const* params, float const* frames,
                             float const* wo, float const* xi, float* out, void* stream);

/* ---- path-guiding network (no counterpart in the reference) ----
 *
 * A small MLP maps a shading point (position in the scene box, normal and
 * outgoing direction, each mapped to [0,1]; 9 inputs) through a
 * 12-frequency sin/cos encoding (216 features, padded to 224) and
 * 224-64-64-64-32 layers (ReLU) to 8 von Mises-Fisher lobes x (a, b, c, d):
 * mean direction (pi sigmoid(a), 2 pi sigmoid(b)) in spherical angles,
 * kappa = exp(clamp(c, -4, 8)), lobe weight softmax(d).  BF16 weights, FP32
 * accumulation and FP32 master parameters trained with Adam on the loss
 * -sum_i (t_i / sum t) log q(w_i).  Parameters (24800 floats): per layer
 * W [in][out] row-major then b [out].  The nearest reference code are the
 * particle-importance maps (core/rendering/integrator/particle/
 * particle_importance.hpp); nothing in sprout learns. */

#define MI_GUIDE_NUM_PARAMS 24800u

int32_t mi_guide_create(uint64_t seed, uint32_t* guide);
int32_t mi_guide_release(uint32_t guide);
/* HOST arrays of MI_GUIDE_NUM_PARAMS floats; set also resets Adam's moments */
int32_t mi_guide_set_params(uint32_t guide, float const* params);
int32_t mi_guide_get_params(uint32_t guide, float* params);
/* The input encoding: feature pair (i, j) of input i is sin / cos of
 * pi (freq[i][j] x_i) (fp32 product), 9 x 12 HOST floats, finite, |f| <= 4096.
 * Default: f = 2^(j/2) for the three position inputs, 2^(j/4) for the normal
 * and direction inputs, j = 0..11.  No counterpart in sprout (see above). */
#define MI_GUIDE_NUM_FREQ 108u
int32_t mi_guide_set_frequencies(uint32_t guide, float const* freq);
/* DEVICE arrays: x n x 9 inputs, out n x 32 raw outputs */
int32_t mi_guide_forward(uint32_t guide, float const* x, uint32_t n, float* out, void* stream);
/* DEVICE arrays: x n x 9, dirs n x 3 unit directions, weights n (>= 0);
 * grad receives MI_GUIDE_NUM_PARAMS floats of the loss gradient */
int32_t mi_guide_gradients(uint32_t guide, float const* x, float const* dirs,
                           float const* weights, uint32_t n, float* grad, void* stream);
/* one Adam step (beta 0.9 / 0.99, eps 1e-8) on the batch */
int32_t mi_guide_train_step(uint32_t guide, float const* x, float const* dirs,
                            float const* weights, uint32_t n, float learning_rate,
                            void* stream);

/* Guided rendering (no counterpart in the reference; its plain estimator is
 * mi_render).  At every non-emissive vertex the network's mixture for
 * (position, normal, wo) and the BSDF form a one-sample mixture: the guide is
 * sampled with probability fraction, the BSDF otherwise, and the path
 * weight uses the mixture pdf fraction * q + (1 - fraction) * p_bsdf, as do
 * the MIS weights against light sampling, so the estimator stays unbiased
 * (fraction 0 is plain MIS).  With train != 0 each vertex of depth below
 * record_depth records (inputs, sampled direction, radiance gathered after
 * it / throughput before it), and after every wave train_steps Adam steps
 * run over those records in slices of batch (0: all records over the
 * steps). */
typedef struct mi_guide_settings {
    uint32_t guide;
    float    fraction;
    uint32_t train;
    uint32_t train_steps;
    uint32_t batch;
    float    learning_rate;
    uint32_t record_depth;
    uint32_t guide_depth; /* vertices of depth below this sample the guide (0: all); deeper
                             ones sample the BSDF alone and record nothing */
    uint32_t medium;      /* nonzero: scattering vertices in the scene's medium are guided
                             and recorded too, the guide mixed with the phase function; their
                             network inputs are the scattering point, a zero normal and wo */
} mi_guide_settings;

/* As mi_render, guided.  With training, every wave is followed on the
 * device by the training weights of its records, a shuffled copy of them
 * and train_steps Adam steps over slices of `batch` records (0: all records
 * in train_steps slices); the whole wave, training included, is one hipGraph
 * (kernel by kernel with MI_RENDER_NO_GRAPH). */
int32_t mi_render_guided(uint32_t scene, mi_camera const* camera,
                         mi_render_settings const* settings, mi_guide_settings const* guide,
                         double* accum, void* stream, mi_render_stats* stats);

/* ---- synthetic scenes (host-side, deterministic; used by tests and bench) ---- */

/* The procedural box room: five inward-facing Lambertian quads of the box
 * [-1,1] x [0,2] x [-1,1] (floor, ceiling, back wall grey 0.75; left wall
 * red 0.63/0.06/0.05; right wall green 0.14/0.45/0.09; the +z side open),
 * two axis-aligned grey boxes, and a downward-facing ceiling quad light of
 * radiance 17/12/4 (x, z in [-0.25, 0.25], y = 1.98).  With slit != 0 the
 * light is enclosed by a box of occluding walls, open only through a slit
 * of that width on its -z side.  Material 0 grey, 1 red, 2 green, 3 light.
 * Pass NULL buffers to query the counts.  HOST pointers. */
int32_t mi_gen_room(float slit, float* positions, uint32_t* num_vertices, uint32_t* indices,
                    uint32_t* num_triangles, uint32_t* triangle_materials,
                    mi_material* materials, uint32_t* num_materials);


/* Volume scene: a grey ground quad (y = 0, 8 x 8) and above it, facing
 * down at y = 3.5, an emissive quad of 6 x 6 split into 1000 triangles
 * (radiance 4 / 4.5 / 5); material 0 ground (Lambert 0.5), 1 light.  The
 * medium is given separately (mi_scene_set_medium).  Pass NULL buffers to
 * query the counts.  HOST pointers. */
int32_t mi_gen_cloud_scene(float* positions, uint32_t* num_vertices, uint32_t* indices,
                           uint32_t* num_triangles, uint32_t* triangle_materials,
                           mi_material* materials, uint32_t* num_materials);

/* Density grid nx x ny x nz (x fastest, HOST) from seeded fBm of `octaves`
 * octaves of value noise inside a spherical envelope, scaled to [0,
 * max_density]. */
int32_t mi_gen_fbm_density(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t octaves, uint64_t seed,
                           float max_density, float* density);

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
