// extern "C" entry points of the sprout_gpu library (include/sprout_gpu.h).

#include "bvh_build.hpp"
#include "common.hpp"
#include "guide.hpp"
#include "instances.hpp"
#include "light_build.hip.hpp"
#include "light_tree.hpp"
#include "render.hpp"
#include "trace.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

namespace mi {

namespace {

thread_local char last_error[512] = "";

struct Context {
    int32_t                   device = -1;
    TraceResources            trace;
    std::vector<DeviceBvh*>   bvhs;    // handle = index
    std::vector<RenderScene*> scenes;  // handle = index
    std::vector<GuideNet*>    guides;  // handle = index
    std::vector<Tlas*>        tlases;  // handle = index
    TraceResources            render_trace;  // the renderer's own ray counter and spill area
    RenderState               render_state;
    float*                    e_table = nullptr;  // GGX energy table, made on first use
    Light_forest_scratch      light_scratch;       // the device light-tree builder's
};

Context* context = nullptr;

hipStream_t as_stream(void* s) {
    return static_cast<hipStream_t>(s);
}

DeviceBvh* lookup(uint32_t handle) {
    if (!context || handle >= context->bvhs.size()) {
        return nullptr;
    }
    return context->bvhs[handle];
}

uint32_t store(DeviceBvh const& bvh) {
    auto* b = new DeviceBvh(bvh);
    for (uint32_t i = 0; i < context->bvhs.size(); ++i) {
        if (!context->bvhs[i]) {
            context->bvhs[i] = b;
            return i;
        }
    }
    context->bvhs.push_back(b);
    return uint32_t(context->bvhs.size() - 1);
}

RenderScene* lookup_scene(uint32_t handle) {
    if (!context || handle >= context->scenes.size()) {
        return nullptr;
    }
    return context->scenes[handle];
}

Tlas* lookup_tlas(uint32_t handle) {
    if (!context || handle >= context->tlases.size()) {
        return nullptr;
    }
    return context->tlases[handle];
}

GuideNet* lookup_guide(uint32_t handle) {
    if (!context || handle >= context->guides.size()) {
        return nullptr;
    }
    return context->guides[handle];
}

int32_t energy_table(float const** table) {
    if (!context->e_table) {
        if (int32_t const st = make_energy_table(&context->e_table); st) {
            return st;
        }
    }
    *table = context->e_table;
    return MI_OK;
}

}  // namespace

int32_t fail(int32_t status, char const* fmt, ...) {
    va_list args;
    va_start(args, fmt);
    std::vsnprintf(last_error, sizeof(last_error), fmt, args);
    va_end(args);
    return status;
}

void clear_error() {
    last_error[0] = '\0';
}

}  // namespace mi

using namespace mi;

extern "C" {

int32_t mi_init(int32_t device) {
    clear_error();
    if (context) {
        return context->device == device ? MI_OK
                                         : fail(MI_ERR_ARGS, "mi_init: already on device %d",
                                                context->device);
    }

    int count = 0;
    if (hipError_t const e = hipGetDeviceCount(&count); e != hipSuccess || device < 0 ||
                                                        device >= count) {
        return fail(MI_ERR_HIP, "mi_init: device %d not available (%d devices)", device, count);
    }
    if (hipError_t const e = hipSetDevice(device); e != hipSuccess) {
        return fail(MI_ERR_HIP, "mi_init: %s", hipGetErrorString(e));
    }

    hipDeviceProp_t prop;
    if (hipError_t const e = hipGetDeviceProperties(&prop, device); e != hipSuccess) {
        return fail(MI_ERR_HIP, "mi_init: %s", hipGetErrorString(e));
    }

    // the fixed scratch of mi_trace_closest_noalloc on this (current) device
    if (int32_t const st = noalloc_scratch_setup(device); st) {
        return st;
    }

    uint32_t* counter = nullptr;
    if (hipError_t const e = hipMalloc(&counter, Counter_words * sizeof(uint32_t)); e != hipSuccess) {
        return fail(MI_ERR_HIP, "mi_init: %s", hipGetErrorString(e));
    }

    context                = new Context;
    context->device        = device;
    context->trace.num_cus = uint32_t(prop.multiProcessorCount);
    context->trace.counter = counter;
    return MI_OK;
}

int32_t mi_release(void) {
    clear_error();
    if (!context) {
        return fail(MI_ERR_NO_CONTEXT, "mi_release: no context");
    }
    (void)hipDeviceSynchronize();
    for (DeviceBvh* b : context->bvhs) {
        if (b) {
            b->release();
            delete b;
        }
    }
    for (RenderScene* sc : context->scenes) {
        if (sc) {
            sc->release();
            delete sc;
        }
    }
    for (GuideNet* g : context->guides) {
        if (g) {
            g->release();
            delete g;
        }
    }
    for (Tlas* t : context->tlases) {
        if (t) {
            t->release();
            delete t;
        }
    }
    context->render_state.release();
    context->render_trace.release();
    context->light_scratch.release();
    (void)hipFree(context->e_table);
    context->trace.release();
    delete context;
    context = nullptr;
    return MI_OK;
}

char const* mi_last_error(void) {
    return last_error;
}

int32_t mi_bvh_build(float const* positions, uint32_t num_vertices, uint32_t position_stride,
                     uint32_t const* indices, uint32_t num_triangles,
                     mi_bvh_settings const* settings, void* stream, uint32_t* bvh) {
    clear_error();
    if (!context) {
        return fail(MI_ERR_NO_CONTEXT, "mi_bvh_build: call mi_init first");
    }
    if (!positions || !indices || !bvh || position_stride < 3 || 0 == num_triangles ||
        0 == num_vertices) {
        return fail(MI_ERR_ARGS, "mi_bvh_build: bad arguments");
    }

    mi_bvh_settings const s = settings ? *settings : mi_bvh_settings{16, 64, 4};
    if (0 == s.num_slices || 0 == s.max_primitives || s.max_primitives > 0xFF) {
        return fail(MI_ERR_ARGS, "mi_bvh_build: bad settings");
    }

    DeviceBvh out;
    if (int32_t const st = build_bvh(positions, num_vertices, position_stride, indices,
                                     num_triangles, s, as_stream(stream), out);
        st != MI_OK) {
        return st;
    }
    *bvh = store(out);
    return MI_OK;
}

int32_t mi_bvh_upload(void const* nodes, uint32_t num_nodes, uint32_t const* slots,
                      uint32_t num_slots, float const* positions, uint32_t num_vertices,
                      uint32_t position_stride, uint32_t const* indices, uint32_t num_triangles,
                      uint32_t* bvh) {
    clear_error();
    if (!context) {
        return fail(MI_ERR_NO_CONTEXT, "mi_bvh_upload: call mi_init first");
    }
    if (!nodes || 0 == num_nodes || (num_slots && !slots) || !positions || !indices || !bvh ||
        position_stride < 3) {
        return fail(MI_ERR_ARGS, "mi_bvh_upload: bad arguments");
    }

    DeviceBvh out;
    if (int32_t const st = upload_bvh(nodes, num_nodes, slots, num_slots, positions, num_vertices,
                                      position_stride, indices, num_triangles, out);
        st != MI_OK) {
        return st;
    }
    *bvh = store(out);
    return MI_OK;
}

int32_t mi_bvh_release(uint32_t bvh) {
    clear_error();
    DeviceBvh* b = lookup(bvh);
    if (!b) {
        return fail(context ? MI_ERR_HANDLE : MI_ERR_NO_CONTEXT, "mi_bvh_release: bad handle %u",
                    bvh);
    }
    for (uint32_t i = 0; i < context->tlases.size(); ++i) {
        Tlas const* t = context->tlases[i];
        if (t && std::find(t->mesh_handles.begin(), t->mesh_handles.end(), bvh) != t->mesh_handles.end()) {
            return fail(MI_ERR_ARGS, "mi_bvh_release: bvh %u is placed by instances %u", bvh, i);
        }
    }
    (void)hipDeviceSynchronize();
    b->release();
    delete b;
    context->bvhs[bvh] = nullptr;
    return MI_OK;
}

int32_t mi_bvh_get_info(uint32_t bvh, mi_bvh_info* info) {
    clear_error();
    DeviceBvh const* b = lookup(bvh);
    if (!b || !info) {
        return fail(context ? MI_ERR_HANDLE : MI_ERR_NO_CONTEXT, "mi_bvh_get_info: bad handle %u",
                    bvh);
    }
    info->num_nodes     = b->num_nodes;
    info->num_slots     = b->num_slots;
    info->num_triangles = b->num_triangles;
    info->max_depth     = b->max_depth;
    for (int k = 0; k < 3; ++k) {
        info->aabb_min[k] = b->aabb_min[k];
        info->aabb_max[k] = b->aabb_max[k];
    }
    info->build_ms = b->build_ms;
    return MI_OK;
}

int32_t mi_bvh_download(uint32_t bvh, void* nodes, uint32_t* slots) {
    clear_error();
    DeviceBvh const* b = lookup(bvh);
    if (!b) {
        return fail(context ? MI_ERR_HANDLE : MI_ERR_NO_CONTEXT, "mi_bvh_download: bad handle %u",
                    bvh);
    }
    hipError_t e = hipSuccess;
    if (nodes) {
        e = hipMemcpy(nodes, b->nodes, size_t(b->num_nodes) * 32, hipMemcpyDeviceToHost);
    }
    if (e == hipSuccess && slots && b->num_slots) {
        e = hipMemcpy(slots, b->slots, size_t(b->num_slots) * 4, hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) {
        return fail(MI_ERR_HIP, "mi_bvh_download: %s", hipGetErrorString(e));
    }
    return MI_OK;
}

int32_t mi_trace_closest(uint32_t bvh, mi_ray const* rays, uint32_t num_rays, mi_hit* hits,
                         void* stream) {
    clear_error();
    DeviceBvh const* b = lookup(bvh);
    if (!b) {
        return fail(context ? MI_ERR_HANDLE : MI_ERR_NO_CONTEXT, "mi_trace_closest: bad handle %u",
                    bvh);
    }
    if (num_rays && (!rays || !hits)) {
        return fail(MI_ERR_ARGS, "mi_trace_closest: null buffers");
    }
    return launch_trace(*b, context->trace, rays, num_rays, hits, nullptr, false,
                        as_stream(stream));
}

int32_t mi_trace_closest_noalloc(uint32_t bvh, mi_ray const* rays, uint32_t num_rays, mi_hit* hits,
                                 void* stream) {
    clear_error();
    DeviceBvh const* b = lookup(bvh);
    if (!b) {
        return fail(context ? MI_ERR_HANDLE : MI_ERR_NO_CONTEXT,
                    "mi_trace_closest_noalloc: bad handle %u", bvh);
    }
    if (num_rays && (!rays || !hits)) {
        return fail(MI_ERR_ARGS, "mi_trace_closest_noalloc: null buffers");
    }
    return launch_trace_noalloc(*b, context->trace, context->device, rays, num_rays, hits,
                                as_stream(stream));
}

int32_t mi_trace_any(uint32_t bvh, mi_ray const* rays, uint32_t num_rays, uint32_t* occluded,
                     void* stream) {
    clear_error();
    DeviceBvh const* b = lookup(bvh);
    if (!b) {
        return fail(context ? MI_ERR_HANDLE : MI_ERR_NO_CONTEXT, "mi_trace_any: bad handle %u",
                    bvh);
    }
    if (num_rays && (!rays || !occluded)) {
        return fail(MI_ERR_ARGS, "mi_trace_any: null buffers");
    }
    return launch_trace(*b, context->trace, rays, num_rays, occluded, nullptr, true,
                        as_stream(stream));
}

int32_t mi_trace_counts(uint32_t bvh, mi_ray const* rays, uint32_t num_rays, int32_t any_hit,
                        void* hits_or_occluded, uint32_t* counts, void* stream) {
    clear_error();
    DeviceBvh const* b = lookup(bvh);
    if (!b) {
        return fail(context ? MI_ERR_HANDLE : MI_ERR_NO_CONTEXT, "mi_trace_counts: bad handle %u",
                    bvh);
    }
    if (num_rays && (!rays || !hits_or_occluded || !counts)) {
        return fail(MI_ERR_ARGS, "mi_trace_counts: null buffers");
    }
    return launch_trace(*b, context->trace, rays, num_rays, hits_or_occluded, counts, any_hit != 0,
                        as_stream(stream));
}

int32_t mi_set_trace_config(mi_trace_config const* config) {
    clear_error();
    if (!context) {
        return fail(MI_ERR_NO_CONTEXT, "mi_set_trace_config: call mi_init first");
    }
    if (!config) {
        return fail(MI_ERR_ARGS, "mi_set_trace_config: null config");
    }
    TraceResources const defaults;
    uint32_t const block = config->block_size ? config->block_size : defaults.block_size;
    uint32_t const bpc   = config->blocks_per_cu ? config->blocks_per_cu : defaults.blocks_per_cu;
    uint32_t const lanes = config->refill_lanes ? config->refill_lanes : defaults.refill_lanes;
    uint32_t const batch = config->claim_batch ? config->claim_batch : defaults.claim_batch;
    if (block % 64 || block > 1024 || bpc > 64 || lanes > 64 || config->kernel > 4 ||
        batch > 65536 || config->lds_entries > 32) {
        return fail(MI_ERR_ARGS,
                    "mi_set_trace_config: block %u, blocks per CU %u, kernel %u, refill lanes %u, "
                    "claim batch %u, lds entries %u",
                    block, bpc, config->kernel, lanes, batch, config->lds_entries);
    }
    context->trace.block_size    = block;
    context->trace.blocks_per_cu = bpc;
    context->trace.kernel        = config->kernel;
    context->trace.refill_lanes  = lanes;
    context->trace.claim_batch   = batch;
    context->trace.lds_entries   = config->lds_entries ? config->lds_entries : defaults.lds_entries;
    context->trace.sort_rays     = config->sort_rays ? 1u : 0u;
    return MI_OK;
}

int32_t mi_trace_autotune(uint32_t bvh, mi_ray const* rays, uint32_t num_rays, uint32_t repeats,
                          void* stream, mi_trace_config* best, float* best_ms) {
    clear_error();
    DeviceBvh const* b = lookup(bvh);
    if (!b) {
        return fail(context ? MI_ERR_HANDLE : MI_ERR_NO_CONTEXT, "mi_trace_autotune: bad handle %u",
                    bvh);
    }
    if (0 == num_rays || !rays || 0 == repeats || repeats > 1000) {
        return fail(MI_ERR_ARGS, "mi_trace_autotune: bad arguments");
    }
    hipStream_t const s    = as_stream(stream);
    mi_hit*           hits = nullptr;
    hipEvent_t        e0 = nullptr, e1 = nullptr;
    hipError_t        e  = hipMalloc(&hits, size_t(num_rays) * sizeof(mi_hit));
    if (!e) e = hipEventCreate(&e0);
    if (!e) e = hipEventCreate(&e1);
    if (e) {
        (void)hipFree(hits);
        return fail(MI_ERR_HIP, "mi_trace_autotune: %s", hipGetErrorString(e));
    }
    // candidate shapes around the measured defaults (block 256, claim 64)
    struct Shape {
        uint32_t lds, lanes, batch, per_cu;
    };
    static Shape const shapes[] = {{10, 8, 64, 8},  {8, 8, 64, 8},   {12, 8, 64, 8},  {10, 4, 64, 8},
                                   {10, 16, 64, 8}, {10, 24, 64, 8}, {10, 8, 32, 8},  {10, 8, 128, 8},
                                   {10, 8, 64, 6},  {10, 8, 64, 10}, {16, 8, 64, 5},  {9, 8, 64, 8}};
    // only the shape fields change; the counter and spill buffers stay the
    // context's own (a launch may grow the spill area)
    auto set_shape = [](TraceResources& t, Shape const& sh, uint32_t kernel, uint32_t block) {
        t.kernel        = kernel;
        t.lds_entries   = sh.lds;
        t.refill_lanes  = sh.lanes;
        t.claim_batch   = sh.batch;
        t.blocks_per_cu = sh.per_cu;
        t.block_size    = block;
    };
    Shape const    old_shape{context->trace.lds_entries, context->trace.refill_lanes,
                          context->trace.claim_batch, context->trace.blocks_per_cu};
    uint32_t const old_kernel = context->trace.kernel;
    uint32_t const old_block  = context->trace.block_size;
    float          best_t      = 0.f;
    Shape          best_shape  = shapes[0];
    uint32_t       best_kernel = 3;
    int32_t        st          = MI_OK;
    for (uint32_t const kernel : {3u, 4u})
    for (Shape const& sh : shapes) {
        TraceResources& t = context->trace;
        set_shape(t, sh, kernel, 256);
        t.sort_rays = 0;
        st = launch_trace(*b, t, rays, num_rays, hits, nullptr, false, s);  // warm-up
        if (!st) e = hipEventRecord(e0, s);
        for (uint32_t k = 0; k < repeats && !st && !e; ++k) {
            st = launch_trace(*b, t, rays, num_rays, hits, nullptr, false, s);
        }
        if (!st && !e) e = hipEventRecord(e1, s);
        if (!st && !e) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (!st && !e) e = hipEventElapsedTime(&ms, e0, e1);
        if (st || e) {
            break;
        }
        ms /= float(repeats);
        if (0.f == best_t || ms < best_t) {
            best_t      = ms;
            best_shape  = sh;
            best_kernel = kernel;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(s);
    (void)hipFree(hits);
    if (st || e) {
        set_shape(context->trace, old_shape, old_kernel, old_block);
        return st ? st : fail(MI_ERR_HIP, "mi_trace_autotune: %s", hipGetErrorString(e));
    }
    TraceResources& t = context->trace;
    set_shape(t, best_shape, best_kernel, 256);
    // the best shape once more with the rays sorted first (the sort counts)
    {
        hipEvent_t f0 = nullptr, f1 = nullptr;
        hipError_t fe = hipEventCreate(&f0);
        if (!fe) fe = hipEventCreate(&f1);
        mi_hit* sh = nullptr;
        if (!fe) fe = hipMalloc(&sh, size_t(num_rays) * sizeof(mi_hit));
        t.sort_rays = 1;
        int32_t fs  = fe ? MI_ERR_HIP : launch_trace(*b, t, rays, num_rays, sh, nullptr, false, s);
        if (!fs && !fe) fe = hipEventRecord(f0, s);
        for (uint32_t k = 0; k < repeats && !fs && !fe; ++k) {
            fs = launch_trace(*b, t, rays, num_rays, sh, nullptr, false, s);
        }
        if (!fs && !fe) fe = hipEventRecord(f1, s);
        if (!fs && !fe) fe = hipEventSynchronize(f1);
        float ms = 0.f;
        if (!fs && !fe) fe = hipEventElapsedTime(&ms, f0, f1);
        t.sort_rays = (!fs && !fe && ms / float(repeats) < best_t) ? 1u : 0u;
        if (t.sort_rays) {
            best_t = ms / float(repeats);
        }
        (void)hipEventDestroy(f0);
        (void)hipEventDestroy(f1);
        (void)hipStreamSynchronize(s);
        (void)hipFree(sh);
    }
    if (best) {
        *best = mi_trace_config{t.block_size, t.blocks_per_cu, t.kernel, t.refill_lanes,
                                t.claim_batch, t.lds_entries, t.sort_rays};
    }
    if (best_ms) {
        *best_ms = best_t;
    }
    return MI_OK;
}

int32_t mi_sync(void* stream) {
    clear_error();
    if (!context) {
        return fail(MI_ERR_NO_CONTEXT, "mi_sync: no context");
    }
    if (hipError_t const e = hipStreamSynchronize(as_stream(stream)); e != hipSuccess) {
        return fail(MI_ERR_HIP, "mi_sync: %s", hipGetErrorString(e));
    }
    return MI_OK;
}

int32_t mi_scene_create(float const* positions, uint32_t num_vertices, uint32_t const* indices,
                        uint32_t num_triangles, uint32_t const* triangle_materials,
                        mi_material const* materials, uint32_t num_materials,
                        mi_bvh_settings const* bvh_settings, uint32_t* scene) {
    clear_error();
    if (!context) {
        return fail(MI_ERR_NO_CONTEXT, "mi_scene_create: call mi_init first");
    }
    if (!positions || !indices || !triangle_materials || !materials || !scene ||
        0 == num_vertices || 0 == num_triangles || 0 == num_materials) {
        return fail(MI_ERR_ARGS, "mi_scene_create: bad arguments");
    }
    mi_bvh_settings const st = bvh_settings ? *bvh_settings : mi_bvh_settings{16, 64, 4};
    if (0 == st.num_slices || 0 == st.max_primitives || st.max_primitives > 0xFF) {
        return fail(MI_ERR_ARGS, "mi_scene_create: bad tree settings");
    }
    RenderScene out;
    if (int32_t const r = create_scene(positions, num_vertices, indices, num_triangles,
                                       triangle_materials, materials, num_materials, st, out);
        r) {
        return r;
    }
    auto* sc = new RenderScene(out);
    for (uint32_t i = 0; i < context->scenes.size(); ++i) {
        if (!context->scenes[i]) {
            context->scenes[i] = sc;
            *scene             = i;
            return MI_OK;
        }
    }
    context->scenes.push_back(sc);
    *scene = uint32_t(context->scenes.size() - 1);
    return MI_OK;
}

int32_t mi_scene_create_instanced(mi_mesh const* meshes, uint32_t num_meshes,
                                  uint32_t const* instance_mesh, float const* world_to_object,
                                  float const* object_to_world,
                                  uint32_t const* instance_materials, uint32_t num_instances,
                                  float const* positions, uint32_t num_vertices,
                                  uint32_t const* indices, uint32_t num_triangles,
                                  uint32_t const* triangle_materials, mi_material const* materials,
                                  uint32_t num_materials, mi_bvh_settings const* bvh_settings,
                                  uint32_t* scene) {
    clear_error();
    if (!context) {
        return fail(MI_ERR_NO_CONTEXT, "mi_scene_create_instanced: call mi_init first");
    }
    if (!meshes || 0 == num_meshes || !instance_mesh || !world_to_object || !object_to_world ||
        !instance_materials || 0 == num_instances || !positions || !indices ||
        !triangle_materials || !materials || !scene || 0 == num_vertices || 0 == num_triangles ||
        0 == num_materials) {
        return fail(MI_ERR_ARGS, "mi_scene_create_instanced: bad arguments");
    }
    mi_bvh_settings const st = bvh_settings ? *bvh_settings : mi_bvh_settings{16, 64, 4};
    if (0 == st.num_slices || 0 == st.max_primitives || st.max_primitives > 0xFF) {
        return fail(MI_ERR_ARGS, "mi_scene_create_instanced: bad tree settings");
    }
    RenderScene out;
    if (int32_t const r = create_scene_instanced(meshes, num_meshes, instance_mesh, world_to_object,
                                                 object_to_world, instance_materials, num_instances,
                                                 positions, num_vertices, indices, num_triangles,
                                                 triangle_materials, materials, num_materials, st,
                                                 out);
        r) {
        return r;
    }
    auto* sc = new RenderScene(out);
    for (uint32_t i = 0; i < context->scenes.size(); ++i) {
        if (!context->scenes[i]) {
            context->scenes[i] = sc;
            *scene             = i;
            return MI_OK;
        }
    }
    context->scenes.push_back(sc);
    *scene = uint32_t(context->scenes.size() - 1);
    return MI_OK;
}

int32_t mi_scene_update_instances(uint32_t scene, float const* world_to_object,
                                  float const* object_to_world, uint32_t num_instances) {
    clear_error();
    RenderScene* sc = lookup_scene(scene);
    if (!sc) {
        return fail(context ? MI_ERR_HANDLE : MI_ERR_NO_CONTEXT,
                    "mi_scene_update_instances: bad handle %u", scene);
    }
    if (!world_to_object || !object_to_world || 0 == num_instances) {
        return fail(MI_ERR_ARGS, "mi_scene_update_instances: bad arguments");
    }
    (void)hipDeviceSynchronize();  // renders in flight read the old top level
    return update_instances(*sc, world_to_object, object_to_world, num_instances);
}

int32_t mi_scene_update_positions(uint32_t scene, float const* positions, uint32_t num_vertices,
                                  double* light_ms, void* stream) {
    clear_error();
    RenderScene* sc = lookup_scene(scene);
    if (!sc) {
        return fail(context ? MI_ERR_HANDLE : MI_ERR_NO_CONTEXT,
                    "mi_scene_update_positions: bad handle %u", scene);
    }
    if (!positions) {
        return fail(MI_ERR_ARGS, "mi_scene_update_positions: null positions");
    }
    (void)hipDeviceSynchronize();  // renders in flight read the old geometry and lights
    return update_positions(*sc, positions, num_vertices, context->light_scratch,
                            as_stream(stream), light_ms);
}

int32_t mi_scene_release(uint32_t scene) {
    clear_error();
    RenderScene* sc = lookup_scene(scene);
    if (!sc) {
        return fail(context ? MI_ERR_HANDLE : MI_ERR_NO_CONTEXT,
                    "mi_scene_release: bad handle %u", scene);
    }
    (void)hipDeviceSynchronize();
    sc->release();
    delete sc;
    context->scenes[scene] = nullptr;
    return MI_OK;
}

namespace {

int32_t render_call(uint32_t scene, mi_camera const* camera, mi_render_settings const* settings,
                    mi_guide_settings const* gs, double* accum, void* stream,
                    mi_render_stats* stats, Sensor_filter const* sensor = nullptr) {
    if (!context) {
        return fail(MI_ERR_NO_CONTEXT, "mi_render: call mi_init first");
    }
    GuideNet* net = nullptr;
    if (gs) {
        net = lookup_guide(gs->guide);
        if (!net) {
            return fail(MI_ERR_HANDLE, "mi_render_guided: bad guide handle %u", gs->guide);
        }
    }
    RenderScene* sc = lookup_scene(scene);
    if (!sc) {
        return fail(MI_ERR_HANDLE, "mi_render: bad scene handle %u", scene);
    }
    if (!camera || !settings || !accum) {
        return fail(MI_ERR_ARGS, "mi_render: bad arguments");
    }
    float const* table = nullptr;
    if (int32_t const st = energy_table(&table); st) {
        return st;
    }
    TraceResources& rt = context->render_trace;
    if (!rt.counter) {
        if (hipError_t const e = hipMalloc(&rt.counter, Counter_words * sizeof(uint32_t)); e) {
            return fail(MI_ERR_HIP, "mi_render: %s", hipGetErrorString(e));
        }
    }
    // launch shape of the traversal as configured by mi_set_trace_config
    rt.block_size    = context->trace.block_size;
    rt.blocks_per_cu = context->trace.blocks_per_cu;
    rt.num_cus       = context->trace.num_cus;
    rt.kernel        = context->trace.kernel;
    rt.refill_lanes  = context->trace.refill_lanes;
    rt.claim_batch   = context->trace.claim_batch;
    rt.lds_entries   = context->trace.lds_entries;
    return render(*sc, rt, table, context->render_state, *camera, *settings, net, gs, accum,
                  as_stream(stream), stats, sensor);
}

}  // namespace

int32_t mi_render_filtered(uint32_t scene, mi_camera const* camera,
                           mi_render_settings const* settings, mi_sensor const* sensor,
                           double* accum, void* stream, mi_render_stats* stats) {
    clear_error();
    if (!sensor) {
        return fail(MI_ERR_ARGS, "mi_render_filtered: null sensor");
    }
    Sensor_filter f;
    if (int32_t const st = make_sensor_filter(*sensor, f); st) {
        return st;
    }
    return render_call(scene, camera, settings, nullptr, accum, stream, stats, &f);
}

int32_t mi_render(uint32_t scene, mi_camera const* camera, mi_render_settings const* settings,
                  double* accum, void* stream, mi_render_stats* stats) {
    clear_error();
    return render_call(scene, camera, settings, nullptr, accum, stream, stats);
}

int32_t mi_render_guided(uint32_t scene, mi_camera const* camera,
                         mi_render_settings const* settings, mi_guide_settings const* guide,
                         double* accum, void* stream, mi_render_stats* stats) {
    clear_error();
    if (!guide) {
        return fail(MI_ERR_ARGS, "mi_render_guided: no guide settings");
    }
    return render_call(scene, camera, settings, guide, accum, stream, stats);
}

int32_t mi_ggx_e_table(float* table, void* stream) {
    clear_error();
    if (!context) {
        return fail(MI_ERR_NO_CONTEXT, "mi_ggx_e_table: call mi_init first");
    }
    if (!table) {
        return fail(MI_ERR_ARGS, "mi_ggx_e_table: bad arguments");
    }
    float const* t = nullptr;
    if (int32_t const st = energy_table(&t); st) {
        return st;
    }
    if (hipError_t const e = hipMemcpyAsync(table, t, 32 * 32 * sizeof(float),
                                            hipMemcpyDeviceToDevice, as_stream(stream));
        e) {
        return fail(MI_ERR_HIP, "mi_ggx_e_table: %s", hipGetErrorString(e));
    }
    return MI_OK;
}

namespace {

int32_t substitute_batch(bool sample, uint32_t n, float const* params, float const* frames,
                         float const* wo, float const* w_or_xi, float* out, void* stream) {
    clear_error();
    if (!context) {
        return fail(MI_ERR_NO_CONTEXT, "mi_substitute_*: call mi_init first");
    }
    if (n && (!params || !frames || !wo || !w_or_xi || !out)) {
        return fail(MI_ERR_ARGS, "mi_substitute_*: bad arguments");
    }
    float const* t = nullptr;
    if (int32_t const st = energy_table(&t); st) {
        return st;
    }
    return bsdf_batch(sample, n, params, frames, wo, w_or_xi, t, out, as_stream(stream));
}

}  // namespace

int32_t mi_substitute_evaluate(uint32_t n, float const* params, float const* frames,
                               float const* wo, float const* wi, float* out, void* stream) {
    return substitute_batch(false, n, params, frames, wo, wi, out, stream);
}

int32_t mi_substitute_sample(uint32_t n, float const* params, float const* frames,
                             float const* wo, float const* xi, float* out, void* stream) {
    return substitute_batch(true, n, params, frames, wo, xi, out, stream);
}

int32_t mi_guide_create(uint64_t seed, uint32_t* guide) {
    clear_error();
    if (!context) {
        return fail(MI_ERR_NO_CONTEXT, "mi_guide_create: call mi_init first");
    }
    if (!guide) {
        return fail(MI_ERR_ARGS, "mi_guide_create: bad arguments");
    }
    GuideNet n;
    if (int32_t const st = guide_create(seed, context->trace.num_cus, n); st) {
        return st;
    }
    auto* g = new GuideNet(n);
    for (uint32_t i = 0; i < context->guides.size(); ++i) {
        if (!context->guides[i]) {
            context->guides[i] = g;
            *guide             = i;
            return MI_OK;
        }
    }
    context->guides.push_back(g);
    *guide = uint32_t(context->guides.size() - 1);
    return MI_OK;
}

int32_t mi_guide_release(uint32_t guide) {
    clear_error();
    GuideNet* g = lookup_guide(guide);
    if (!g) {
        return fail(context ? MI_ERR_HANDLE : MI_ERR_NO_CONTEXT, "mi_guide_release: bad handle %u",
                    guide);
    }
    (void)hipDeviceSynchronize();
    g->release();
    delete g;
    context->guides[guide] = nullptr;
    return MI_OK;
}

int32_t mi_guide_set_params(uint32_t guide, float const* params) {
    clear_error();
    GuideNet* g = lookup_guide(guide);
    if (!g || !params) {
        return fail(g ? MI_ERR_ARGS : MI_ERR_HANDLE, "mi_guide_set_params: bad arguments");
    }
    return guide_set_params(*g, params);
}

int32_t mi_guide_set_frequencies(uint32_t guide, float const* freq) {
    clear_error();
    GuideNet* g = lookup_guide(guide);
    if (!g || !freq) {
        return fail(g ? MI_ERR_ARGS : MI_ERR_HANDLE, "mi_guide_set_frequencies: bad arguments");
    }
    (void)hipDeviceSynchronize();
    return guide_set_frequencies(*g, freq);
}

int32_t mi_guide_get_params(uint32_t guide, float* params) {
    clear_error();
    GuideNet* g = lookup_guide(guide);
    if (!g || !params) {
        return fail(g ? MI_ERR_ARGS : MI_ERR_HANDLE, "mi_guide_get_params: bad arguments");
    }
    (void)hipDeviceSynchronize();
    return guide_get_params(*g, params);
}

int32_t mi_guide_forward(uint32_t guide, float const* x, uint32_t n, float* out, void* stream) {
    clear_error();
    GuideNet* g = lookup_guide(guide);
    if (!g || (n && (!x || !out))) {
        return fail(g ? MI_ERR_ARGS : MI_ERR_HANDLE, "mi_guide_forward: bad arguments");
    }
    return guide_forward(*g, x, n, out, as_stream(stream));
}

int32_t mi_guide_gradients(uint32_t guide, float const* x, float const* dirs,
                           float const* weights, uint32_t n, float* grad, void* stream) {
    clear_error();
    GuideNet* g = lookup_guide(guide);
    if (!g || 0 == n || !x || !dirs || !weights || !grad) {
        return fail(g ? MI_ERR_ARGS : MI_ERR_HANDLE, "mi_guide_gradients: bad arguments");
    }
    if (int32_t const st = guide_gradients(*g, x, dirs, weights, n, as_stream(stream)); st) {
        return st;
    }
    if (hipError_t const e = hipMemcpyAsync(grad, g->grad, guide::Num_params * sizeof(float),
                                            hipMemcpyDeviceToDevice, as_stream(stream));
        e) {
        return fail(MI_ERR_HIP, "mi_guide_gradients: %s", hipGetErrorString(e));
    }
    return MI_OK;
}

int32_t mi_guide_train_step(uint32_t guide, float const* x, float const* dirs,
                            float const* weights, uint32_t n, float learning_rate,
                            void* stream) {
    clear_error();
    GuideNet* g = lookup_guide(guide);
    if (!g || 0 == n || !x || !dirs || !weights || !(learning_rate > 0.f)) {
        return fail(g ? MI_ERR_ARGS : MI_ERR_HANDLE, "mi_guide_train_step: bad arguments");
    }
    return guide_adam_step(*g, x, dirs, weights, n, learning_rate, as_stream(stream));
}

int32_t mi_light_tree_build(float const* box_min, float const* box_max, float const* cones,
                            float const* powers, uint8_t const* two_sided, uint32_t num_lights,
                            mi_light_node* nodes, uint32_t* middles, uint32_t* mapping,
                            uint32_t* orders, uint32_t* num_nodes) {
    clear_error();
    if (!num_nodes || (num_lights && (!box_min || !box_max || !cones || !powers || !two_sided))) {
        return fail(MI_ERR_ARGS, "mi_light_tree_build: bad arguments");
    }
    HostLightTree t;
    build_light_tree(box_min, box_max, cones, powers, two_sided, num_lights, t);
    *num_nodes = uint32_t(t.nodes.size());
    if (nodes) {
        if (!middles || !mapping || !orders) {
            return fail(MI_ERR_ARGS, "mi_light_tree_build: bad arguments");
        }
        std::memcpy(nodes, t.nodes.data(), t.nodes.size() * sizeof(mi_light_node));
        std::memcpy(middles, t.middles.data(), t.middles.size() * sizeof(uint32_t));
        std::memcpy(mapping, t.light_mapping.data(), t.light_mapping.size() * sizeof(uint32_t));
        std::memcpy(orders, t.light_orders.data(), t.light_orders.size() * sizeof(uint32_t));
    }
    return MI_OK;
}

int32_t mi_light_tree_build_device(float const* box_min, float const* box_max, float const* cones,
                                   float const* powers, uint8_t const* two_sided,
                                   uint32_t num_lights, mi_light_node* nodes, uint32_t* middles,
                                   uint32_t* mapping, uint32_t* orders, uint32_t* num_nodes,
                                   void* stream) {
    clear_error();
    if (!context) {
        return fail(MI_ERR_NO_CONTEXT, "mi_light_tree_build_device: call mi_init first");
    }
    if (!num_nodes || (num_lights && (!box_min || !box_max || !cones || !powers || !two_sided ||
                                      !nodes || !middles || !mapping || !orders))) {
        return fail(MI_ERR_ARGS, "mi_light_tree_build_device: bad arguments");
    }
    *num_nodes = 0;
    if (0 == num_lights) {
        return MI_OK;
    }
    uint32_t const   first[2] = {0, num_lights};
    Light_forest     in;
    in.set        = {box_min, box_max, cones, powers, two_sided};
    in.first      = first;
    in.num_roots  = 1;
    in.num_lights = num_lights;
    Light_forest_out out;
    out.nodes          = nodes;
    out.middles        = middles;
    out.mapping        = mapping;
    out.orders         = orders;
    out.num_nodes_host = num_nodes;
    int32_t const st   = build_light_forest(in, context->light_scratch, out, as_stream(stream));
    if (MI_ERR_UNSPLITTABLE == st) {
        return fail(st, "mi_light_tree_build_device: a node no candidate splits (build on the host)");
    }
    if (st) {
        return fail(st, "mi_light_tree_build_device: %s", hipGetErrorString(hipGetLastError()));
    }
    return MI_OK;
}

int32_t mi_light_part_trees_build_device(float const* box_min, float const* box_max,
                                         float const* cones, float const* powers,
                                         uint32_t num_lights, uint32_t const* first,
                                         uint32_t num_parts, float const* part_cones,
                                         mi_light_node* nodes, uint32_t* middles,
                                         uint32_t* mapping, uint32_t* orders,
                                         uint32_t* cdf_offset, float* cdfs, uint32_t* num_nodes,
                                         uint32_t* num_cdfs, void* stream) {
    clear_error();
    if (!context) {
        return fail(MI_ERR_NO_CONTEXT, "mi_light_part_trees_build_device: call mi_init first");
    }
    if (!first || (num_parts && (!box_min || !box_max || !cones || !powers || !part_cones ||
                                 !nodes || !middles || !mapping || !orders || !cdf_offset ||
                                 !cdfs || !num_nodes || !num_cdfs))) {
        return fail(MI_ERR_ARGS, "mi_light_part_trees_build_device: bad arguments");
    }
    if (0 == num_parts) {
        return MI_OK;
    }
    if (0 != first[0] || first[num_parts] != num_lights) {
        return fail(MI_ERR_ARGS, "mi_light_part_trees_build_device: first must run from 0 to %u",
                    num_lights);
    }
    for (uint32_t r = 0; r < num_parts; ++r) {
        if (first[r + 1] <= first[r]) {
            return fail(MI_ERR_ARGS, "mi_light_part_trees_build_device: part %u is empty", r);
        }
    }
    Light_forest in;
    in.set        = {box_min, box_max, cones, powers, nullptr};
    in.first      = first;
    in.part_cones = part_cones;
    in.num_roots  = num_parts;
    in.num_lights = num_lights;
    in.part       = true;
    Light_forest_out out;
    out.nodes          = nodes;
    out.middles        = middles;
    out.mapping        = mapping;
    out.orders         = orders;
    out.cdf_offset     = cdf_offset;
    out.cdfs           = cdfs;
    out.num_cdfs       = num_cdfs;
    out.num_nodes_host = num_nodes;
    if (int32_t const st = build_light_forest(in, context->light_scratch, out, as_stream(stream)); st) {
        return fail(st, "mi_light_part_trees_build_device: %s", hipGetErrorString(hipGetLastError()));
    }
    return MI_OK;
}

int32_t mi_light_part_tree_build(float const* box_min, float const* box_max, float const* cones,
                                 float const* powers, uint32_t num_lights, float const* part_cone,
                                 mi_light_node* nodes, uint32_t* middles, uint32_t* mapping,
                                 uint32_t* orders, uint32_t* cdf_offset, float* cdfs,
                                 uint32_t* num_nodes, uint32_t* num_cdfs) {
    clear_error();
    if (!num_nodes || !num_cdfs || !part_cone ||
        (num_lights && (!box_min || !box_max || !cones || !powers))) {
        return fail(MI_ERR_ARGS, "mi_light_part_tree_build: bad arguments");
    }
    HostLightTree t;
    build_part_tree(box_min, box_max, cones, powers, num_lights, part_cone, t);
    *num_nodes = uint32_t(t.nodes.size());
    *num_cdfs  = uint32_t(t.cdfs.size());
    if (nodes) {
        if (!middles || !mapping || !orders || !cdf_offset || (!cdfs && !t.cdfs.empty())) {
            return fail(MI_ERR_ARGS, "mi_light_part_tree_build: bad arguments");
        }
        std::memcpy(nodes, t.nodes.data(), t.nodes.size() * sizeof(mi_light_node));
        std::memcpy(middles, t.middles.data(), t.middles.size() * sizeof(uint32_t));
        std::memcpy(mapping, t.light_mapping.data(), t.light_mapping.size() * sizeof(uint32_t));
        std::memcpy(orders, t.light_orders.data(), t.light_orders.size() * sizeof(uint32_t));
        std::memcpy(cdf_offset, t.cdf_offset.data(), t.cdf_offset.size() * sizeof(uint32_t));
        if (!t.cdfs.empty()) {
            std::memcpy(cdfs, t.cdfs.data(), t.cdfs.size() * sizeof(float));
        }
    }
    return MI_OK;
}

int32_t mi_light_tree_sample(uint32_t scene, float const* points, float const* normals,
                             float const* randoms, uint32_t split, uint32_t n, uint32_t* picks,
                             uint32_t* counts, void* stream) {
    clear_error();
    RenderScene* sc = lookup_scene(scene);
    if (!sc || (n && (!points || !normals || !randoms || !picks || !counts))) {
        return fail(sc ? MI_ERR_ARGS : MI_ERR_HANDLE, "mi_light_tree_sample: bad arguments");
    }
    return light_tree_queries(*sc, true, points, normals, randoms, split, n, picks, counts,
                              as_stream(stream));
}

int32_t mi_light_tree_pdf(uint32_t scene, float const* points, float const* normals,
                          uint32_t const* lights, uint32_t split, uint32_t n, float* pdfs,
                          void* stream) {
    clear_error();
    RenderScene* sc = lookup_scene(scene);
    if (!sc || (n && (!points || !normals || !lights || !pdfs))) {
        return fail(sc ? MI_ERR_ARGS : MI_ERR_HANDLE, "mi_light_tree_pdf: bad arguments");
    }
    return light_tree_queries(*sc, false, points, normals, lights, split, n, pdfs, nullptr,
                              as_stream(stream));
}

int32_t mi_instances_create(uint32_t const* meshes, float const* world_to_object,
                            float const* object_to_world, uint32_t num_instances,
                            mi_bvh_settings const* settings, void* stream, uint32_t* instances) {
    clear_error();
    if (!context) {
        return fail(MI_ERR_NO_CONTEXT, "mi_instances_create: call mi_init first");
    }
    if (!meshes || !world_to_object || !object_to_world || !instances || 0 == num_instances) {
        return fail(MI_ERR_ARGS, "mi_instances_create: bad arguments");
    }
    mi_bvh_settings const st = settings ? *settings : mi_bvh_settings{16, 64, 4};
    // distinct mesh trees, in order of first use
    std::vector<uint32_t>         handles;
    std::vector<DeviceBvh const*> trees;
    std::vector<uint32_t>         mesh_of(num_instances);
    for (uint32_t i = 0; i < num_instances; ++i) {
        DeviceBvh const* b = lookup(meshes[i]);
        if (!b) {
            return fail(MI_ERR_HANDLE, "mi_instances_create: instance %u: bad bvh handle %u", i,
                        meshes[i]);
        }
        auto const it = std::find(handles.begin(), handles.end(), meshes[i]);
        if (it == handles.end()) {
            mesh_of[i] = uint32_t(handles.size());
            handles.push_back(meshes[i]);
            trees.push_back(b);
        } else {
            mesh_of[i] = uint32_t(it - handles.begin());
        }
    }
    auto* t = new Tlas();
    if (int32_t const r = create_tlas(trees.data(), uint32_t(trees.size()), mesh_of.data(),
                                      world_to_object, object_to_world, num_instances, st,
                                      as_stream(stream), *t);
        r) {
        delete t;
        return r;
    }
    t->mesh_handles = handles;
    for (uint32_t i = 0; i < context->tlases.size(); ++i) {
        if (!context->tlases[i]) {
            context->tlases[i] = t;
            *instances         = i;
            return MI_OK;
        }
    }
    context->tlases.push_back(t);
    *instances = uint32_t(context->tlases.size() - 1);
    return MI_OK;
}

int32_t mi_instances_release(uint32_t instances) {
    clear_error();
    Tlas* t = lookup_tlas(instances);
    if (!t) {
        return fail(context ? MI_ERR_HANDLE : MI_ERR_NO_CONTEXT,
                    "mi_instances_release: bad handle %u", instances);
    }
    (void)hipDeviceSynchronize();
    t->release();
    delete t;
    context->tlases[instances] = nullptr;
    return MI_OK;
}

namespace {

int32_t trace_instances_call(uint32_t instances, mi_ray const* rays, uint32_t n, void* out,
                             bool any_hit, void* stream) {
    clear_error();
    Tlas const* t = lookup_tlas(instances);
    if (!t) {
        return fail(context ? MI_ERR_HANDLE : MI_ERR_NO_CONTEXT, "instance trace: bad handle %u",
                    instances);
    }
    if (n && (!rays || !out)) {
        return fail(MI_ERR_ARGS, "instance trace: null buffer");
    }
    return trace_instances_launch(*t, context->trace.num_cus, rays, n, nullptr, out, any_hit,
                                  as_stream(stream));
}

}  // namespace

int32_t mi_trace_instances_closest(uint32_t instances, mi_ray const* rays, uint32_t num_rays,
                                   mi_instance_hit* hits, void* stream) {
    return trace_instances_call(instances, rays, num_rays, hits, false, stream);
}

int32_t mi_trace_instances_any(uint32_t instances, mi_ray const* rays, uint32_t num_rays,
                               uint32_t* occluded, void* stream) {
    return trace_instances_call(instances, rays, num_rays, occluded, true, stream);
}

int32_t mi_instances_download(uint32_t instances, void* nodes, uint32_t* slots, float* boxes,
                              uint32_t* num_nodes, uint32_t* num_slots) {
    clear_error();
    Tlas const* t = lookup_tlas(instances);
    if (!t || !num_nodes || !num_slots) {
        return fail(t ? MI_ERR_ARGS : MI_ERR_HANDLE, "mi_instances_download: bad arguments");
    }
    *num_nodes = t->tree.num_nodes;
    *num_slots = t->tree.num_slots;
    if (boxes) {
        for (uint32_t i = 0; i < t->num_instances; ++i) {
            for (int k = 0; k < 3; ++k) {
                boxes[6 * size_t(i) + k]     = t->records[i].lo[k];
                boxes[6 * size_t(i) + 3 + k] = t->records[i].hi[k];
            }
        }
    }
    hipError_t e = hipSuccess;
    if (nodes) {
        e = hipMemcpy(nodes, t->tree.nodes, size_t(t->tree.num_nodes) * 32, hipMemcpyDeviceToHost);
    }
    if (e == hipSuccess && slots && t->tree.num_slots) {
        e = hipMemcpy(slots, t->tree.slots, size_t(t->tree.num_slots) * 4, hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) {
        return fail(MI_ERR_HIP, "mi_instances_download: %s", hipGetErrorString(e));
    }
    return MI_OK;
}

int32_t mi_scene_set_normals(uint32_t scene, float const* normals, uint32_t num_vertices) {
    clear_error();
    RenderScene* sc = lookup_scene(scene);
    if (!sc) {
        return fail(context ? MI_ERR_HANDLE : MI_ERR_NO_CONTEXT,
                    "mi_scene_set_normals: bad handle %u", scene);
    }
    (void)hipDeviceSynchronize();  // renders in flight may read the old normals
    return set_normals(*sc, normals, num_vertices);
}

int32_t mi_scene_set_mesh_normals(uint32_t scene, uint32_t mesh, float const* normals,
                                  uint32_t num_vertices) {
    clear_error();
    RenderScene* sc = lookup_scene(scene);
    if (!sc) {
        return fail(context ? MI_ERR_HANDLE : MI_ERR_NO_CONTEXT,
                    "mi_scene_set_mesh_normals: bad handle %u", scene);
    }
    (void)hipDeviceSynchronize();  // renders in flight may read the old normals
    return set_mesh_normals(*sc, mesh, normals, num_vertices);
}

int32_t mi_scene_set_medium(uint32_t scene, float const* density, uint32_t const dims[3],
                            float const box_lo[3], float const box_hi[3], float const cc_a[3],
                            float const cc_s[3], float g) {
    clear_error();
    RenderScene* sc = lookup_scene(scene);
    if (!sc) {
        return fail(context ? MI_ERR_HANDLE : MI_ERR_NO_CONTEXT,
                    "mi_scene_set_medium: bad handle %u", scene);
    }
    if (density && (!dims || !box_lo || !box_hi || !cc_a || !cc_s)) {
        return fail(MI_ERR_ARGS, "mi_scene_set_medium: null argument");
    }
    (void)hipDeviceSynchronize();  // renders in flight may read the old medium
    return set_medium(*sc, density, dims, box_lo, box_hi, cc_a, cc_s, g);
}

int32_t mi_scene_set_medium_majorants(uint32_t scene, float const* density, uint32_t const dims[3],
                                      float const box_lo[3], float const box_hi[3],
                                      float const cc_a[3], float const cc_s[3], float g,
                                      uint32_t majorants) {
    clear_error();
    RenderScene* sc = lookup_scene(scene);
    if (!sc) {
        return fail(context ? MI_ERR_HANDLE : MI_ERR_NO_CONTEXT,
                    "mi_scene_set_medium_majorants: bad handle %u", scene);
    }
    if (density && (!dims || !box_lo || !box_hi || !cc_a || !cc_s)) {
        return fail(MI_ERR_ARGS, "mi_scene_set_medium_majorants: null argument");
    }
    (void)hipDeviceSynchronize();  // renders in flight may read the old medium
    return set_medium(*sc, density, dims, box_lo, box_hi, cc_a, cc_s, g, majorants);
}

int32_t mi_medium_gridtree(float const* density, uint32_t const dims[3], float const cc_a[3],
                           float const cc_s[3], uint32_t* num_cells, uint32_t* nodes,
                           uint32_t* num_nodes, float* data, uint32_t* num_data) {
    return mi_medium_majorant_tree(density, dims, cc_a, cc_s, MI_MAJORANTS_GRIDTREE, num_cells,
                                   nodes, num_nodes, data, num_data);
}

int32_t mi_medium_majorant_tree(float const* density, uint32_t const dims[3], float const cc_a[3],
                                float const cc_s[3], uint32_t majorants, uint32_t* num_cells,
                                uint32_t* nodes, uint32_t* num_nodes, float* data,
                                uint32_t* num_data) {
    clear_error();
    if (majorants > MI_MAJORANTS_FLAT) {
        return fail(MI_ERR_ARGS, "mi_medium_majorant_tree: majorants %u unknown", majorants);
    }
    if (!density || !dims || !cc_a || !cc_s || !num_nodes || !num_data) {
        return fail(MI_ERR_ARGS, "mi_medium_majorant_tree: null argument");
    }
    for (int k = 0; k < 3; ++k) {
        if (0 == dims[k] || dims[k] > 4096) {
            return fail(MI_ERR_ARGS, "mi_medium_majorant_tree: dimension %u out of [1, 4096]", dims[k]);
        }
    }
    int32_t const idims[3] = {int32_t(dims[0]), int32_t(dims[1]), int32_t(dims[2])};
    HostGridtree  tree;
    build_gridtree(density, idims, cc_a, cc_s, tree, MI_MAJORANTS_FLAT == majorants);
    if (num_cells) {
        for (int k = 0; k < 3; ++k) {
            num_cells[k] = uint32_t(tree.num_cells[k]);
        }
    }
    if (nodes) {
        if (*num_nodes < tree.nodes.size()) {
            return fail(MI_ERR_ARGS, "mi_medium_majorant_tree: %zu nodes do not fit in %u",
                        tree.nodes.size(), *num_nodes);
        }
        std::memcpy(nodes, tree.nodes.data(), tree.nodes.size() * sizeof(uint32_t));
    }
    if (data) {
        if (*num_data < tree.data.size()) {
            return fail(MI_ERR_ARGS, "mi_medium_gridtree: %zu leaves do not fit in %u",
                        tree.data.size(), *num_data);
        }
        std::memcpy(data, tree.data.data(), tree.data.size() * sizeof(Medium_cm));
    }
    *num_nodes = uint32_t(tree.nodes.size());
    *num_data  = uint32_t(tree.data.size());
    return MI_OK;
}

int32_t mi_medium_track(uint32_t scene, mi_ray const* rays, uint32_t depth, uint64_t seed,
                        uint32_t flags, uint32_t n, float* out, void* stream) {
    clear_error();
    RenderScene* sc = lookup_scene(scene);
    if (!sc) {
        return fail(context ? MI_ERR_HANDLE : MI_ERR_NO_CONTEXT, "mi_medium_track: bad handle %u",
                    scene);
    }
    if (n && (!rays || !out)) {
        return fail(MI_ERR_ARGS, "mi_medium_track: null argument");
    }
    return medium_queries(*sc, rays, depth, seed, flags & (MI_MEDIUM_TRANSMITTANCE | MI_MEDIUM_NEAREST),
                          n, out, as_stream(stream));
}

int32_t mi_scene_light_tree(uint32_t scene, mi_light_node* nodes, uint32_t* middles,
                            uint32_t* mapping, uint32_t* orders, void* infos,
                            uint32_t* num_nodes, uint32_t* num_lights) {
    clear_error();
    RenderScene* sc = lookup_scene(scene);
    if (!sc || !num_nodes || !num_lights) {
        return fail(sc ? MI_ERR_ARGS : MI_ERR_HANDLE, "mi_scene_light_tree: bad arguments");
    }
    *num_nodes  = sc->num_lights ? sc->num_light_nodes : 0;
    *num_lights = sc->num_lights;
    return download_light_tree(*sc, nodes, middles, mapping, orders, infos);
}

}  // extern "C"
