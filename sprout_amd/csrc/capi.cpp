// extern "C" entry points of the sprout_gpu library (include/sprout_gpu.h).

#include "bvh_build.hpp"
#include "common.hpp"
#include "trace.hpp"

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

namespace mi {

namespace {

thread_local char last_error[512] = "";

struct Context {
    int32_t                 device = -1;
    TraceResources          trace;
    std::vector<DeviceBvh*> bvhs;  // handle = index
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

    uint32_t* counter = nullptr;
    if (hipError_t const e = hipMalloc(&counter, sizeof(uint32_t)); e != hipSuccess) {
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
    if (block % 64 || block > 1024 || bpc > 64 || lanes > 64 || config->kernel > 2 ||
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

}  // extern "C"
