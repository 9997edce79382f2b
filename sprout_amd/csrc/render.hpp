// Scene, per-render state and entry points of the wavefront path tracer (host code).
#pragma once

#include "common.hpp"
#include "guide.hpp"
#include "instances.hpp"
#include "light_tree.hpp"
#include "medium.hpp"
#include "trace.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace mi {

// Device material record (48 bytes, read as three float4 by the shading kernel).
struct Render_material {
    float    color[3];
    float    roughness;
    float    emission[3];
    float    metallic;
    float    ior;
    uint32_t kind;  // MI_MATERIAL_*
    uint32_t pad[2];
};

// A sensor filter as the kernels evaluate it: sprout's 30-sample tables
// (Gaussian over d^2, Blackman over |d|) or Mitchell's polynomial, and the clamp.
struct Sensor_filter {
    uint32_t kind;              // mi_filter_kind
    int32_t  border;            // pixels a sample reaches: 0 (box), 1 or 2
    float    table[30];
    float    range_end;
    float    inverse_interval;
    float    radius_inv;
    float    b, c;
    uint32_t clamp_mode;        // 0 none, 1 per channel, 2 luminance
    float    clamp_max[3];
};

// Builds the tables of a sensor as sprout's filter constructors do.
int32_t make_sensor_filter(mi_sensor const& sensor, Sensor_filter& out);

// Vertex data of an instanced mesh on the device.
struct Mesh_geometry {
    float const*    positions;  // 3 floats per vertex
    uint32_t const* indices;    // 3 per triangle
    float const*    normals;    // 3 per vertex: shading normals, or null (geometric)
};

// Device buffers of update_positions: the builders' light inputs.
struct Emitter_work {
    float*    tri_lo    = nullptr;  // 3 per emissive triangle
    float*    tri_hi    = nullptr;
    float*    tri_cone  = nullptr;  // 4 per emissive triangle
    float*    tri_power = nullptr;
    float*    part_cone = nullptr;  // 4 per part
    float*    part_lo   = nullptr;  // 3 per part
    float*    part_hi   = nullptr;
    float*    part_pow  = nullptr;
    uint8_t*  part_two  = nullptr;
    uint32_t* cdf_rel   = nullptr;  // per pool node, relative to its part
    uint32_t* num_cdfs  = nullptr;  // per part
    uint32_t* part_first = nullptr; // per part + 1 (device)
    uint32_t* bad       = nullptr;  // an emissive triangle without area
    uint32_t  capacity_tris  = 0;
    uint32_t  capacity_parts = 0;

    void release();
};

struct RenderScene {
    DeviceBvh        bvh;
    float*           positions    = nullptr;  // 3 floats per vertex
    float*           normals      = nullptr;  // 3 per vertex: shading normals, or null (geometric)
    uint32_t*        indices      = nullptr;  // 3 per triangle
    uint32_t*        tri_material = nullptr;
    Render_material* materials    = nullptr;
    Light_node*      light_nodes   = nullptr;  // light tree over the emissive triangles
    uint32_t*        light_middles = nullptr;
    uint32_t*        light_mapping = nullptr;
    uint32_t*        light_orders  = nullptr;
    Light_info*      light_infos   = nullptr;
    uint32_t*        tri_light     = nullptr;  // emissive part of each triangle, or Null
    uint32_t*        tri_local     = nullptr;  // index of an emissive triangle in its part
    // Primitive_tree pools of the emissive parts (see dev::Part_pools)
    dev::Part_info*  part_infos      = nullptr;
    Light_node*      part_nodes      = nullptr;
    uint32_t*        part_middles    = nullptr;
    uint32_t*        part_cdf_offset = nullptr;
    float*           part_cdfs       = nullptr;
    uint32_t*        part_mapping    = nullptr;
    uint32_t*        part_orders     = nullptr;
    Light_info*      part_lights     = nullptr;
    // what moving the emitters needs: the emissive triangles in part order,
    // each part's first (host), its radiance, and the pools' layout
    uint32_t*                    emit_tris     = nullptr;
    float*                       part_radiance = nullptr;
    std::vector<uint32_t>        part_first;
    std::vector<dev::Part_info>  part_info_host;
    uint32_t                     num_emitters  = 0;
    Emitter_work                 emitter_work;
    // participating medium in an axis-aligned box (none while density is null)
    float*           medium_density = nullptr;  // nx * ny * nz, x fastest
    uint32_t*        medium_nodes   = nullptr;  // majorant gridtree
    Medium_cm*       medium_data    = nullptr;
    dev::Medium_view medium{};
    // instanced scenes: the mesh instances and, last, the triangles above
    // (emitters and other world-space triangles) as one more instance
    Tlas                   tlas;
    std::vector<DeviceBvh> mesh_trees;
    std::vector<void*>     mesh_buffers;             // device positions / indices per mesh
    Mesh_geometry*         mesh_geo         = nullptr;  // per mesh
    std::vector<Mesh_geometry> mesh_geo_host;             // its host copy
    std::vector<uint32_t>  mesh_vertices;                 // vertices per mesh
    std::vector<float*>    mesh_normals;                  // device shading normals per mesh
    float*                 inst_o2w         = nullptr;  // 16 per instance
    uint32_t*              inst_material    = nullptr;  // per instance; Null for the world triangles
    uint32_t               emitter_instance = 0xFFFFFFFFu;
    bool                   instanced        = false;
    mi_bvh_settings        tlas_settings{16, 64, 4};
    mi_bvh_settings        bvh_settings{16, 64, 4};  // of the triangles' tree
    float                  world_box_lo[3] = {0.f, 0.f, 0.f};  // bounds of the world triangles
    float                  world_box_hi[3] = {0.f, 0.f, 0.f};

    uint32_t num_vertices  = 0;
    uint32_t num_triangles = 0;
    uint32_t num_materials = 0;
    uint32_t num_lights      = 0;
    uint32_t num_light_nodes = 0;
    float    box_lo[3]     = {0.f, 0.f, 0.f};  // bounds of all vertices at creation (the
    float    box_hi[3]     = {0.f, 0.f, 0.f};  // guide's position inputs; moves keep them)

    void release();
    void release_medium();
};

struct Light_forest_scratch;

// New positions for the vertices of a (non-instanced) scene, a device array
// of num_vertices x 3 floats with the scene's vertex count: the emissive
// triangles' lights, every part's Primitive_tree and the scene light tree
// are rebuilt on the device (the same trees create_scene builds on the
// host), then the scene's tree over its triangles (the GPU builder with the
// scene's settings).  Fails, keeping the old positions, when an emissive triangle
// would lose its area.  light_ms: the light part's host wall time.  No render
// may be in flight.
int32_t update_positions(RenderScene& s, float const* positions, uint32_t num_vertices,
                         Light_forest_scratch& scratch, hipStream_t stream, double* light_ms);

// An instanced scene: num_instances placements of the meshes (row-vector
// 4x4 matrices, one material each, not emissive) plus the world-space
// triangles of the arrays (emitters among them), which form the light tree
// as in create_scene and are traced as one more instance.  Host arrays.
int32_t create_scene_instanced(mi_mesh const* meshes, uint32_t num_meshes,
                               uint32_t const* instance_mesh, float const* world_to_object,
                               float const* object_to_world, uint32_t const* instance_materials,
                               uint32_t num_instances, float const* positions,
                               uint32_t num_vertices, uint32_t const* indices,
                               uint32_t num_triangles, uint32_t const* tri_material,
                               mi_material const* materials, uint32_t num_materials,
                               mi_bvh_settings const& settings, RenderScene& out);

// New matrices for the instances of an instanced scene (host arrays, one
// pair per instance, as at creation): rebuilds the top-level tree on the
// device and updates the instance transforms; meshes, materials and the
// light tree (whose emitters are world triangles) stay.  On failure the
// scene keeps its old placement.  No render may be in flight.
int32_t update_instances(RenderScene& scene, float const* world_to_object,
                         float const* object_to_world, uint32_t num_instances);

// Shading normals of the scene's triangles: 3 floats per vertex (host),
// normalized here; a hit shades with the normals of its triangle's vertices
// interpolated at the hit and normalized.  normals == nullptr goes back to
// the geometric normal.  Mesh instances keep the geometric normal.
int32_t set_normals(RenderScene& scene, float const* normals, uint32_t num_vertices);

// The same for mesh `mesh` of an instanced scene: its vertex normals in
// object space, moved to world space at a hit as the geometric normal is.
int32_t set_mesh_normals(RenderScene& scene, uint32_t mesh, float const* normals,
                         uint32_t num_vertices);

// Density grid (nx * ny * nz floats, x fastest, host memory) filling the box
// [box_lo, box_hi]; cc_a / cc_s: absorption and scattering coefficients per
// unit density (rgb); g: Henyey-Greenstein anisotropy.  Replaces any medium
// the scene had; density == nullptr removes it.
int32_t set_medium(RenderScene& scene, float const* density, uint32_t const dims[3],
                   float const box_lo[3], float const box_hi[3], float const cc_a[3],
                   float const cc_s[3], float g, uint32_t majorants = MI_MAJORANTS_GRIDTREE);

// Buffers of one render, kept between calls and grown on demand.
struct RenderState {
    uint32_t     capacity    = 0;  // paths
    void*        paths       = nullptr;
    mi_ray*      rays[2]     = {nullptr, nullptr};
    uint32_t*    ray_path[2] = {nullptr, nullptr};
    mi_hit*      hits        = nullptr;
    mi_instance_hit* inst_hits = nullptr;  // closest hits of instanced scenes
    mi_ray*      shadow_rays = nullptr;
    void*        shadows     = nullptr;
    uint32_t*    occluded    = nullptr;
    uint2*       shadow_range = nullptr;
    void*        nee          = nullptr;  // per queue entry: what its shadow rays need
    void*        picks        = nullptr;  // per shadow ray: its light pick
    uint32_t*    medium_event = nullptr;  // per queue entry
    float*       medium_p     = nullptr;  // 3 per queue entry
    uint32_t*    counters    = nullptr;  // queue counters, wave base, ray totals
    hipStream_t  stream      = nullptr;  // internal stream the bounce loop runs (and is captured) on
    hipEvent_t   begin       = nullptr;
    hipEvent_t   end         = nullptr;
    hipGraphExec_t       graph = nullptr;  // the last wave's graph, replayed while its key holds
    std::vector<uint8_t> graph_key;
    // guiding
    uint32_t  guide_capacity  = 0;
    float*    guide_x         = nullptr;  // network inputs per queue entry
    float*    guide_raw       = nullptr;  // network outputs per queue entry
    float*    guide_mix       = nullptr;  // prepared mixtures of guided vertices with light picks
    uint32_t  record_capacity = 0;
    float*    train_x         = nullptr;
    float*    train_dir       = nullptr;
    float*    train_w         = nullptr;
    float*    rec_t           = nullptr;
    float*    rec_r           = nullptr;
    uint32_t* rec_path        = nullptr;
    uint32_t* rec_block       = nullptr;  // valid records per 256 slots, then their offsets
    float2*   pixel_samples   = nullptr;  // Golden_ratio pixel samples of a frame
    uint64_t  pixel_sample_capacity = 0;
    float*    sample_rad      = nullptr;  // regenerating: radiance of every sample of a batch
    uint64_t  sample_capacity = 0;
    uint32_t* done            = nullptr;  // regenerating: slots of the paths that ended this bounce
    uint32_t  done_capacity   = 0;
    uint32_t* live_host       = nullptr;  // pinned: the next queue's length after a graph replay
    hipGraphExec_t       regen_graph = nullptr;  // two regenerating bounces, replayed while its key holds
    std::vector<uint8_t> regen_key;
    float2*   path_spec       = nullptr;  // spectral: per path slot {fourth throughput, hero}
    float*    nee_thr4        = nullptr;  // spectral: per queue entry
    uint32_t  spectral_capacity = 0;
    float*    shadow_c4       = nullptr;  // spectral: per shadow entry
    uint32_t  spectral_shadow_capacity = 0;
    float*    shuf_x          = nullptr;  // the wave's records, shuffled for training
    float*    shuf_dir        = nullptr;
    float*    shuf_w          = nullptr;

    uint32_t  shadow_capacity = 0;

    int32_t reserve(uint32_t num_paths, uint32_t num_shadow);
    int32_t reserve_guiding(uint32_t num_paths, uint32_t records);
    int32_t reserve_pixel_samples(uint64_t entries);
    int32_t reserve_regeneration(uint64_t samples, uint32_t slots);
    int32_t reserve_spectral(uint32_t num_paths, uint32_t num_shadow);
    void    release();
};

int32_t make_energy_table(float** table);

int32_t bsdf_batch(bool sample, uint32_t n, float const* params, float const* frames,
                   float const* wo, float const* wi_or_xi, float const* table, float* out,
                   hipStream_t stream);

// Host arrays in; builds the tree on the device with the given settings.
int32_t create_scene(float const* positions, uint32_t num_vertices, uint32_t const* indices,
                     uint32_t num_triangles, uint32_t const* tri_material,
                     mi_material const* materials, uint32_t num_materials,
                     mi_bvh_settings const& settings, RenderScene& out);

// Batch light-tree queries on device arrays (sample: arg = randoms,
// out0 = picks, out1 = counts; pdf: arg = light ids, out0 = pdfs).
int32_t light_tree_queries(RenderScene const& scene, bool sample, float const* points,
                           float const* normals, void const* arg, uint32_t split, uint32_t n,
                           void* out0, uint32_t* out1, hipStream_t stream);

// Batch medium tracking on device rays (flags: MI_MEDIUM_*); out n x 5 floats.
int32_t medium_queries(RenderScene const& scene, mi_ray const* rays, uint32_t depth, uint64_t seed,
                       uint32_t flags, uint32_t n, float* out, hipStream_t stream);

int32_t download_light_tree(RenderScene const& scene, mi_light_node* nodes, uint32_t* middles,
                            uint32_t* mapping, uint32_t* orders, void* infos);

// net == nullptr: plain MIS; else guided by net with the settings gs.
int32_t render(RenderScene const& scene, TraceResources& trace, float const* e_table,
               RenderState& state, mi_camera const& camera, mi_render_settings const& settings,
               GuideNet* net, mi_guide_settings const* gs, double* accum, hipStream_t stream,
               mi_render_stats* stats, Sensor_filter const* sensor = nullptr);

}  // namespace mi
