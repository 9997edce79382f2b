/*
 * ORACLE (test infrastructure only; never linked into the product).
 * CPU restatement of sprout's triangle SBVH build and traversal, Substitute
 * material and MIS path tracer, used by the tests to check the HIP path and
 * by bench.py as the "port" CPU baseline.
 */
#ifndef SPROUT_ORACLE_H
#define SPROUT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORACLE_ERR_ARGS        = -1,
    ORACLE_ERR_UNSPLITTABLE = -3,
    ORACLE_ERR_STACK       = -4,
    ORACLE_ERR_ALLOC       = -5,
};

/* scene::bvh::Node, core/scene/bvh/scene_bvh_node.hpp:44-59 */
typedef struct {
    float    lo[3];
    uint32_t data; /* first child (inner) or first slot (leaf) */
    float    hi[3];
    uint8_t  axis;
    uint8_t  num; /* 0 = inner node, else number of slots */
    uint8_t  pad[2];
} onode;

typedef struct {
    onode*    nodes;
    uint32_t* slots; /* input triangle id per tree slot */
    uint32_t  num_nodes;
    uint32_t  num_slots;
    uint32_t  spatial_split_threshold;
} oracle_bvh;

/* reciprocal modes for the traversal */
enum {
    ORACLE_RCP_PINNED = 0, /* rcp(x) := correctly rounded 1/x, then the reference's Newton step */
    ORACLE_RCP_NATIVE = 1, /* rcp(x) := this host's rcpss/rcpps, then the reference's Newton step */
};

int32_t oracle_sbvh_build(float const* positions, uint32_t num_vertices, uint32_t const* indices,
                          uint32_t num_triangles, uint32_t num_slices, uint32_t sweep_threshold,
                          uint32_t max_primitives, oracle_bvh* out);

void oracle_bvh_free(oracle_bvh* bvh);

float oracle_surface_area(float const lo[3], float const hi[3]);

/*
 * rays: 8 floats per ray {o.xyz, min_t, d.xyz, max_t}.
 * any == 0: hits = 4 words per ray {t, u, v, input triangle id or 0xFFFFFFFF}
 * any != 0: hits = 1 word per ray (1 occluded / 0 free)
 * counts (optional): 2 words per ray {nodes fetched, triangle tests}, counted
 * as the reference's loop performs them.
 * Returns 0, or ORACLE_ERR_STACK if a ray needed more than 127 stack entries.
 */
int32_t oracle_trace(onode const* nodes, uint32_t const* slots, float const* positions,
                     uint32_t const* indices, float const* rays, uint64_t num_rays, int32_t any,
                     int32_t rcp_mode, uint32_t* hits, uint32_t* counts, uint32_t num_threads);

/* Closest hit by testing every triangle in input order with the same
 * primitive test and tie rule (a later equal-t triangle wins). */
void oracle_brute_force(float const* positions, uint32_t const* indices, uint32_t num_triangles,
                        float const* rays, uint64_t num_rays, int32_t rcp_mode, uint32_t* hits);

/* triangle::intersect (Möller-Trumbore as the reference compiles it):
 * 1 and *max_t, *u, *v updated on a hit in [min_t, *max_t] */
int32_t oracle_triangle(float const* pa, float const* pb, float const* pc, float const o[3],
                        float const d[3], float min_t, float* max_t, float* u, float* v,
                        int32_t rcp_mode);

/* ---- instanced meshes (instances.c) ---- */

typedef struct {
    oracle_bvh const* bvh;
    float const*      positions;
    uint32_t const*   indices;
} oracle_mesh;

typedef struct {
    oracle_bvh         top;   /* over the instances' world boxes */
    float*             boxes; /* 6 per instance: lo, hi */
    float const*       w2o;   /* 16 per instance */
    uint32_t const*    mesh_of;
    oracle_mesh const* meshes;
    uint32_t           num_instances;
} oracle_instances;

void    oracle_transform_point(float const m[16], float const v[3], float out[3]);
void    oracle_transform_vector(float const m[16], float const v[3], float out[3]);
void    oracle_transform_box(float const lo[3], float const hi[3], float const m[16],
                             float out_lo[3], float out_hi[3]);
int32_t oracle_instances_build(oracle_mesh const* meshes, uint32_t num_meshes,
                               uint32_t const* mesh_of, float const* world_to_object,
                               float const* object_to_world, uint32_t num_instances,
                               oracle_instances* out);
void    oracle_instances_free(oracle_instances* t);
/* rays 8 floats each; any: 1 word per ray; closest: 5 words {t, u, v,
 * triangle, instance} (pinned reciprocal) */
int32_t oracle_instances_trace(oracle_instances const* t, float const* rays, uint64_t num_rays,
                               int32_t any, uint32_t* hits);

/* The reciprocals as defined above, exposed for tests. */
float oracle_reciprocal(float x, int32_t rcp_mode);
void  oracle_reciprocal3(float const d[3], int32_t rcp_mode, float out[3]);

/* offset_ray, core/scene/ray_offset.inl:18-29 */
void oracle_offset_ray(float const p[3], float const n[3], float out[3]);

/* ---- Substitute BSDF (bsdf.c) ---- */

/* params per case: color.rgb, roughness, metallic, ior; frames: t.xyz, b.xyz,
 * n.xyz, geo_n.xyz; evaluate out: reflection.rgb (times n.wi), pdf;
 * sample xi: lobe, 2D; sample out: wi.xyz, reflection.rgb, pdf, type bits. */
void oracle_substitute_evaluate(uint32_t n, float const* params, float const* frames,
                                float const* wo, float const* wi, float* out);
void oracle_substitute_sample(uint32_t n, float const* params, float const* frames,
                              float const* wo, float const* xi, float* out);

/* 32 x 32 directional albedo table (row = alpha), regenerated by the
 * reference's procedure. */
void oracle_ggx_e_table(float* out);

float oracle_schlick_f0(float n0, float n1);
void  oracle_orthonormal_basis(float const n[3], float t[3], float b[3]);
void  oracle_disk_concentric(float u, float v, float out[2]);

/* ---- light tree sampling (light_tree.c) ---- */

#define ORACLE_MAX_LIGHTS 64u

/* one light (an emissive triangle): same layout as the library's light records */
typedef struct {
    float    center[3];
    float    radius;
    float    cone[4];
    float    power;
    uint32_t two_sided;
    uint32_t triangle;
    float    area;
} oracle_light;

/* light::Node, core/scene/light/light_tree.hpp:32-65 */
typedef struct {
    float    center[3];
    float    radius;
    float    cone[4];
    float    power;
    float    variance;
    uint32_t bits; /* has_children : 1, two_sided : 1, children_or_light : 30 */
    uint32_t num_lights;
} oracle_light_node;

struct oracle_part;

typedef struct {
    oracle_light_node const* nodes;
    uint32_t const*          middles;
    uint32_t const*          mapping; /* order -> light */
    uint32_t const*          orders;  /* light -> order */
    oracle_light const*      lights;
    uint32_t                 num_nodes;
    uint32_t                 num_lights;
    struct oracle_part const* parts; /* scene trees: the Primitive_tree of each light (a part) */
} oracle_light_tree;

/* The Primitive_tree of one emissive part over its triangles (power = area):
 * light_tree.cpp:777-862, with Distribution_1D leaves of more than 4 lights
 * (distribution_1d.inl:30-44). */
typedef struct oracle_part {
    oracle_light_tree tree;
    uint32_t const*   cdf_offset; /* per node: its distribution in cdfs, or 0xFFFFFFFF */
    float const*      cdfs;
} oracle_part;

void  oracle_part_random(oracle_part const* part, float const p[3], float const n[3], int32_t ts,
                         float random, uint32_t* local, float* pdf);
float oracle_part_pdf(oracle_part const* part, float const p[3], float const n[3], int32_t ts,
                      uint32_t local);
/* Part::init's cone (triangle_mesh.cpp:177-239): the dominant axis of the
 * triangle normals weighted by area and the cosine of the widest angle */
void  oracle_part_cone(float const* normals, float const* areas, uint32_t n, float out[4]);

/* total_sphere: the point is in a medium (translucent), lights behind its
 * normal count fully */
float    oracle_light_importance(float const p[3], float const n[3], float const center[3],
                                 float const cone[4], float radius, float power, int two_sided,
                                 int total_sphere);
uint32_t oracle_light_random(oracle_light_tree const* tree, float const p[3], float const n[3],
                             int32_t total_sphere, float random, int32_t split, uint32_t* lights,
                             float* pdfs);
float    oracle_light_pdf(oracle_light_tree const* tree, float const p[3], float const n[3],
                          int32_t total_sphere, int32_t split, uint32_t light);

/* ---- heterogeneous medium (medium.c) ---- */

/* majorant gridtree: nodes (bit 0 has children, bits 1-31 first child or
 * data index, 0x7FFFFFFF empty), data 4 floats per leaf */
typedef struct {
    uint32_t* nodes;
    float*    data;
    uint32_t  num_nodes;
    uint32_t  num_data;
    int32_t   num_cells[3];
    int32_t   dims[3];
} oracle_gridtree;

int32_t oracle_gridtree_build(float const* density, int32_t const dims[3], float const cc_a[3],
                              float const cc_s[3], oracle_gridtree* out);
/* uniform != 0: every cell split to its 8^3-voxel leaves (flat majorants). */
int32_t oracle_majorant_tree_build(float const* density, int32_t const dims[3], float const cc_a[3],
                                   float const cc_s[3], int32_t uniform, oracle_gridtree* out);
void    oracle_gridtree_free(oracle_gridtree* t);

typedef struct {
    float const*           density; /* dims[0] * dims[1] * dims[2], x fastest */
    int32_t                dims[3];
    oracle_gridtree const* tree;
    float                  box_lo[3];
    float                  box_hi[3];
    float                  cc_a[3];
    float                  cc_s[3];
    float                  g;
} oracle_medium;

/* the next uniform random number in [0, 1) of the caller's generator */
typedef float (*oracle_random_fn)(void* state);

float   oracle_offset_f(float t);
float   oracle_medium_density(oracle_medium const* m, float const uvw[3], int32_t nearest);
float   oracle_similarity_scale(float g, uint32_t depth);
/* [t0, t1]: the part of [min_t, max_t] inside the box; 0 if none */
int32_t oracle_medium_overlap(oracle_medium const* m, float const o[3], float const d[3],
                              float min_t, float max_t, float* t0, float* t1);
/* 0 passed, 1 scattered at *t_out, 2 aborted; w: the tracking weight */
int32_t oracle_medium_track(oracle_medium const* m, float const o[3], float const d[3], float t0,
                            float t1, uint32_t depth, int32_t nearest, oracle_random_fn rnd,
                            void* rs, float* t_out, float w[3]);
/* 1 transmitted (tr), 0 blocked */
int32_t oracle_medium_transmittance(oracle_medium const* m, float const o[3], float const d[3],
                                    float t0, float t1, uint32_t depth, int32_t nearest,
                                    oracle_random_fn rnd, void* rs, float tr[3]);
/* The same over nc (3 or 4) channels with the given coefficients per unit
 * density (a hero-wavelength path's four wavelengths) */
int32_t oracle_medium_track_n(oracle_medium const* m, float const o[3], float const d[3],
                              float t0, float t1, uint32_t depth, int32_t nearest,
                              oracle_random_fn rnd, void* rs, int32_t nc, float const* cc_a,
                              float const* cc_s, float* t_out, float* w);
int32_t oracle_medium_transmittance_n(oracle_medium const* m, float const o[3], float const d[3],
                                      float t0, float t1, uint32_t depth, int32_t nearest,
                                      oracle_random_fn rnd, void* rs, int32_t nc,
                                      float const* cc_a, float const* cc_s, float* tr);
float   oracle_phase_hg(float cos_theta, float g);
void    oracle_sample_phase(float const wo[3], float r0, float r1, float g, float wi[3], float* pdf);

/* PCG32 as the path tracer uses it, for batch checks: start(seed, sequence) */
typedef struct {
    uint64_t state;
    uint64_t inc;
} oracle_pcg;
void     oracle_pcg_start(oracle_pcg* g, uint64_t seed, uint64_t sequence);
float    oracle_pcg_float(void* g);
uint32_t oracle_pcg_uint(oracle_pcg* g);

/* ---- camera samples and the perspective camera (sampler.c) ---- */

/* The generator sequence of pixel (x, y) of a frame: rendering_worker.cpp:97-104
 * over the border-extended grid of grid_width x grid_height pixels. */
uint64_t oracle_pixel_sequence(int32_t x, int32_t y, uint32_t grid_width, uint32_t grid_height,
                               int32_t border, uint32_t iteration);

/* The spp camera samples of one pixel from sprout's Golden_ratio sampler: the
 * generator started at (0, sequence), then image (2 spp floats), lens (2 spp)
 * and time (spp) dimensions as Sampler::camera_sample generates them on the
 * first sample.  lens and time may be NULL. */
int32_t oracle_golden_ratio_camera_samples(uint32_t spp, uint64_t sequence, float* image,
                                           float* lens, float* time);

typedef struct {
    float left_top[3];
    float d_x[3];
    float d_y[3];
} oracle_perspective;

/* Perspective::on_update for a horizontal field of view in degrees (no lens tilt or shift) */
void oracle_perspective_setup(uint32_t width, uint32_t height, float fov_deg,
                              oracle_perspective* cam);

/* Perspective::generate_ray of pixel (x, y) with the pixel sample (u, v), for a
 * camera at position with rotation rows r0, r1, r2 (9 floats, unit scale). */
void oracle_perspective_ray(oracle_perspective const* cam, float const position[3],
                            float const rotation[9], int32_t x, int32_t y, float u, float v,
                            float origin[3], float direction[3]);

/* Batch form of oracle_medium_track / _transmittance over n rays (8 floats
 * each: origin, min_t, direction, max_t), generator (seed, index) per ray;
 * out n x 5 as the library's mi_medium_track. */
int32_t oracle_medium_batch(oracle_medium const* m, float const* rays, uint32_t n, uint32_t depth,
                            uint64_t seed, int32_t transmittance, int32_t nearest, float* out);

/* ---- MIS path tracer (pt.c) ---- */

/* same layout as mi_material */
typedef struct {
    float    color[3];
    float    emission[3];
    float    roughness;
    float    metallic;
    float    ior;
    uint32_t kind; /* 0 Substitute, 1 Lambert, 2 light */
} oracle_material;

/* sprout's Perspective camera (no lens): position, rotation rows r0 (right),
 * r1 (up), r2 (view direction), horizontal field of view in degrees. */
typedef struct {
    uint32_t width;
    uint32_t height;
    float    position[3];
    float    rotation[9];
    float    fov_deg;
} oracle_camera;

enum { ORACLE_PT_MIS = 0, ORACLE_PT_BSDF = 1, ORACLE_PT_NEE = 2 };

/* ---- sensor filters (sensor.c) ---- */

enum {
    ORACLE_FILTER_BOX      = 0,
    ORACLE_FILTER_GAUSSIAN = 1,
    ORACLE_FILTER_BLACKMAN = 2,
    ORACLE_FILTER_MITCHELL = 3
};

typedef struct {
    int32_t kind;
    float   radius;
    float   alpha;
    float   b, c;
    float   table[30]; /* Gaussian over d^2 in [0, r^2], Blackman over |d| in [0, r] */
    float   range_end;
    float   inverse_interval;
    float   radius_inv;
} oracle_filter;

void  oracle_filter_init(oracle_filter* f, int32_t kind, float radius, float alpha, float b,
                         float c);
float oracle_filter_eval(oracle_filter const* f, float d);
void  oracle_filter_eval_many(oracle_filter const* f, float const* d, uint32_t n, float* out);

/* An instanced scene for oracle_render: tlas places the meshes and, as its
 * instance `world` (identity matrices), the tree over the render's triangles,
 * whose materials and lights apply to it; every other instance shows its
 * mesh with the one material material[instance]. */
typedef struct {
    oracle_instances const* tlas;
    float const*            o2w;      /* 16 per instance */
    uint32_t const*         material; /* per instance */
    uint32_t                world;
    float const* const*     normals;  /* per mesh: shading normals per vertex, or NULL
                                         (the array itself may be NULL) */
} oracle_pt_instances;

typedef struct {
    uint32_t samples_per_pixel;
    uint32_t first_sample;
    uint32_t min_bounces;
    uint32_t max_bounces;
    uint64_t seed;
    int32_t  mode; /* ORACLE_PT_* */
    uint32_t num_threads;
    int32_t  split_lights; /* split the light tree below depth 3 (sprout's Adaptive) */
    oracle_medium const* medium; /* NULL: none */
    oracle_pt_instances const* instances; /* NULL: the triangles only */
    uint32_t border;  /* pixels rendered beyond each image edge (a filter's radius) */
    int32_t  golden_ratio; /* pixel samples from the Golden_ratio sampler (first_sample 0) */
    float*   samples; /* non-NULL: per sample and grid pixel r, g, b, jitter x, y of the
                         (W + 2 border) x (H + 2 border) grid, sample-major; accum unused */
    int32_t  spectral; /* hero-wavelength transport over four wavelengths (spectral.c) */
    float const* normals; /* per vertex of the triangles, 3 each: the shading normals,
                             interpolated at a hit; NULL: the geometric normal */
} oracle_pt_settings;

/* ---- hero-wavelength transport (spectral.c) ---- */

/* the hero's uniform number of the path with generator sequence `sequence` */
float oracle_spectral_hero(uint64_t seed, uint64_t sequence);
/* wavelength k (0-3) of a path with hero number u, in nm */
float oracle_spectral_wavelength(float u, int32_t k);
/* band weights (red, green, blue) at lambda nm; they sum to one */
void  oracle_spectral_bands(float lambda, float b[3]);
/* an rgb value at the wavelength with band weights b */
float oracle_spectral_upsample(float const b[3], float const x[3]);
/* the inverse of the bands' Gram matrix over the uniform wavelength measure */
void  oracle_spectral_response(float minv[9]);
/* rgb of values at four wavelengths with band weights b */
void  oracle_spectral_to_rgb(float const minv[9], float const b[4][3], float const v[4],
                             float rgb[3]);

/* Adds each pixel's radiance sum over the samples into accum (W * H * 3), or
 * with set->samples writes every sample of the border-extended grid there.
 * tree: the light tree over the scene's emissive triangles. */
int32_t oracle_render(float const* positions, uint32_t num_vertices, uint32_t const* indices,
                      uint32_t num_triangles, uint32_t const* tri_material,
                      oracle_material const* materials, uint32_t num_materials,
                      oracle_light_tree const* tree, oracle_camera const* cam,
                      oracle_pt_settings const* set, double* accum);

#ifdef __cplusplus
}
#endif

#endif
