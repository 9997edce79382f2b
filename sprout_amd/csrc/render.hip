// Wavefront MIS path tracer (sprout's Pathtracer_MIS) on gfx950.
//
//   Pathtracer_MIS::integrate        core/rendering/integrator/surface/pathtracer_mis.cpp:95-279
//   sample_lights / evaluate_light   pathtracer_mis.cpp:281-347
//   connect_light                    pathtracer_mis.cpp:349-380
//   power_heuristic, predivided_power_heuristic, russian_roulette
//                                    core/rendering/integrator/integrator_helper.hpp:32-52
//   Mesh::sample / Mesh::pdf (area light, solid angle pdf d^2 / (cos * area))
//                                    core/scene/shape/triangle/triangle_mesh.cpp:462-562
//   Intersection::offset_p, evaluate_radiance
//                                    core/scene/prop/prop_intersection.inl:67-121
//   offset_ray                       core/scene/ray_offset.inl:18-29
//   rnd::Generator (PCG32)           base/random/generator.inl:16-49
//
// A wave is samples_per_wave samples of every pixel, one path per (sample
// plane, pixel).  Per bounce:
//   trace   closest hit of the live rays (the BVH kernel; the ray count is
//           read on the device)
//   shade   one thread per live path: emission at the hit (MIS-weighted past
//           the first vertex), the depth limit, Russian roulette, next-event
//           estimation (one light sample; the shadow ray is queued with its
//           contribution), BSDF sample, next ray queued.  Queues are
//           compacted with a wave ballot and one atomic per wave.
//   shadow  any hit of the queued shadow rays; unoccluded contributions are
//           added to their path's radiance.
// Every path owns its radiance, so no two threads add into one word, and at
// the end of a wave the planes are summed per pixel in plane order into the
// caller's FP64 accumulator: results do not depend on scheduling.
//
// Random numbers of a path, in this order (shared with the CPU restatement
// that checks this file): 2 for the pixel jitter; per vertex 1 for Russian
// roulette when it applies, 3 for the light sample (light choice, point on
// the triangle) when the scene has lights, 3 for the BSDF sample (lobe
// choice, 2-D direction sample).
//
// The bounce loop has a fixed shape (max_bounces + 1 iterations whose kernels
// read their counts on the device and whose sample base lives in device
// memory), so one wave is captured into a hipGraph and replayed per wave.

#include "bsdf.hpp"
#include "bvh_build.hpp"
#include "common.hpp"
#include "device_math.hpp"
#include "guide.hpp"
#include "light_build.hip.hpp"
#include "light_build.hpp"
#include "light_tree.hpp"
#include "medium.hpp"
#include "spectral.hpp"
#include "render.hpp"
#include "trace.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

namespace mi {

namespace {

using dev::V3;
using dev::v3;

constexpr float Ray_max_t = 3.40282347e+38f;

// Counter words of RenderState::counters, each on its own 256-byte line:
// atomics on one line serialize (about 12 ns each), so the two queues the
// shading kernel appends to must not share one.
constexpr uint32_t Counter_line = 64;
constexpr uint32_t C_queue0     = 0;                 // live rays, two ping-pong queues (lines 0, 1)
constexpr uint32_t C_shadow     = 2 * Counter_line;  // shadow rays
constexpr uint32_t C_base       = 3 * Counter_line;  // first sample index of the current wave
constexpr uint32_t C_records    = 4 * Counter_line;  // training samples recorded in the current wave
constexpr uint32_t C_totals     = 5 * Counter_line;  // two uint64: extension and shadow rays traced,
                                                     // then one uint64: guiding records trained on
constexpr uint32_t C_done       = 6 * Counter_line;  // regenerating: paths that ended this bounce
constexpr uint32_t C_next       = 7 * Counter_line;  // regenerating: the next sample of the batch to start
constexpr uint32_t Num_counters = 8 * Counter_line;

// rnd::Generator, base/random/generator.inl:16-49
struct Rng {
    uint64_t state;
    uint64_t inc;

    __device__ void start(uint64_t s, uint64_t sequence) {
        state = 0;
        inc   = (sequence << 1u) | 1u;
        next();
        state += s;
        next();
    }

    __device__ uint32_t next() {
        uint64_t const old = state;
        state              = old * 6364136223846793005ull + (inc | 1);
        uint32_t const xrs = uint32_t(((old >> 18u) ^ old) >> 27u);
        uint32_t const rot = uint32_t(old >> 59u);
        return (xrs >> rot) | (xrs << ((0u - rot) & 31u));
    }

    __device__ float uniform() {
        uint32_t const bits = (next() & 0x007FFFFFu) | 0x3F800000u;
        return __uint_as_float(bits) - 1.f;
    }

    // The state n steps ahead in O(log n) steps: the n-fold composition of
    // the LCG (Brown, "Random Number Generation with Arbitrary Strides", 1994)
    __device__ void advance(uint64_t n) {
        uint64_t mul     = 6364136223846793005ull;
        uint64_t add     = inc | 1;
        uint64_t acc_mul = 1;
        uint64_t acc_add = 0;
        while (n) {
            if (n & 1) {
                acc_mul *= mul;
                acc_add = acc_add * mul + add;
            }
            add = (mul + 1) * add;
            mul *= mul;
            n >>= 1;
        }
        state = acc_mul * state + acc_add;
    }
};

constexpr uint32_t Flag_singular    = 1;  // State::Treat_as_singular
constexpr uint32_t Flag_translucent = 2;  // State::Is_translucent: the vertex was in a medium
constexpr uint32_t Flag_primary     = 4;  // State::Primary_ray: no non-specular bounce yet
// the medium's density is looked up without filtering for the shadow rays of
// the current vertex and the ray it spawned: Filter::Nearest once the
// vertex's depth is above 1 and the path is no longer primary
// (pathtracer_mis.cpp:121)
constexpr uint32_t Flag_nearest     = 8;

// medium_event values per queue entry
constexpr uint32_t Medium_none     = 0;
constexpr uint32_t Medium_scatter  = 1;
constexpr uint32_t Medium_absorbed = 2;

// Pathtracer_MIS::Num_dedicated_samplers: light-tree splitting below this depth
constexpr uint32_t Num_split_bounces = 3;

struct Path {
    float    throughput[3];
    float    radiance[3];
    float    bsdf_pdf;  // pdf of the direction sample that spawned the current ray
    uint32_t flags;
    Rng      rng;
    uint32_t record;     // training sample of the previous vertex, or Null
    float    prev_n[3];  // geometric normal of the previous vertex
    float    jitter[2];  // the camera sample's position in its pixel
    uint32_t depth;      // depth of the vertex the path's ray leads to (0: camera ray)
    uint32_t sample;     // the sample of the frame batch: plane * num_pixels + pixel
};

struct Shadow {
    float    contrib[3];
    uint32_t path;
};

// What the shadow rays of a vertex need (one per queue entry with picks).
struct Nee_vertex {
    uint64_t rng_state;  // the path's generator after the light-selection draw
    uint64_t rng_inc;
    float    p[3];       // origin of the shadow rays
    uint32_t path;
    float    n[3];       // geometric normal; (0, 0, 1) in the medium
    uint32_t material;   // the surface's material, or Kind_phase
    float    ns[3];      // shading normal (the material's frame)
    float    wo[3];
    uint32_t flags;      // Nee_guided, Nee_medium
    float    throughput[3];
    uint32_t depth;      // the vertex's depth
};

constexpr uint32_t Nee_guided = 1;  // the guide's mixture (guide_mix) is in the MIS weight
constexpr uint32_t Nee_medium = 2;  // a scattering vertex of the medium

// One light pick in its shadow-queue slot: pick k of a queue entry's vertex.
struct Pick_slot {
    uint32_t light;  // the part, or Null: no ray
    float    pdf;
    uint32_t vertex;
    uint32_t k;
};

struct Wave {
    // scene
    float const*           positions;
    float const*           normals;  // shading normals per vertex, or null
    uint32_t const*        indices;
    uint32_t const*        tri_material;
    Render_material const* materials;
    dev::Light_tree_view   tree;
    uint32_t const*        tri_light;  // light of an emissive triangle
    uint32_t               split_lights;
    dev::Part_pools        parts;      // the Primitive_tree of every emissive part
    uint32_t const*        tri_local;  // an emissive triangle's index in its part
    dev::Medium_view       medium;
    // instanced scenes
    // camera grid: the image grown by border pixels on every side (filtered sensor)
    uint32_t               grid_w;
    int32_t                border;
    Sensor_filter          filter;
    uint32_t               instanced;
    uint32_t               emitter_instance;  // the instance of the world triangles
    mi_instance_hit const* ihits;
    Tlas_view              tlas;
    Mesh_geometry const*   mesh_geo;
    float const*           inst_o2w;
    uint32_t const*        inst_material;
    uint32_t*              medium_event;  // per queue entry: Medium_*
    float*                 medium_p;      // scattering point per queue entry
    float const*           e_table;
    // camera: Perspective's position, rotation rows and on_update frame
    float         position[3];
    float         rotation[9];  // rows r0 (right), r1 (up), r2 (view direction)
    float         left_top[3];
    float         d_x[3];
    float         d_y[3];
    uint32_t      width;
    uint32_t      height;
    float2 const* pixel_samples;  // Golden_ratio: table_spp samples per grid pixel, or null
    uint32_t      table_spp;
    // settings
    uint32_t min_bounces;
    uint32_t max_bounces;
    uint64_t seed;
    uint32_t num_paths;
    uint32_t num_pixels;
    uint32_t batch_samples;  // samples of the frame batch (a wave's num_paths without regeneration)
    uint32_t regenerate;     // ended paths hand their slot to the batch's next sample
    float*   sample_rad;     // regenerating: radiance of every sample of the batch (3 floats)
    uint32_t* done;          // regenerating: slots of the paths that ended this bounce
    // state
    Path*     paths;
    mi_ray*   rays[2];
    uint32_t* ray_path[2];
    mi_hit*   hits;
    mi_ray*   shadow_rays;
    Shadow*   shadows;
    uint32_t* occluded;
    uint2*    shadow_range;  // first shadow entry and count of each queue entry's vertex
    Nee_vertex* nee;         // per queue entry whose vertex picked lights
    Pick_slot*  picks;       // per shadow entry: the light it goes to
    uint32_t* counters;
    // guiding
    float        fraction;      // probability of sampling the guide
    float const* guide_raw;     // 32 network outputs per queue entry
    float*       guide_mix;     // the mixture of a guided vertex with light picks (Vmf_mixture::store)
    float*       guide_x;       // 9 network inputs per queue entry
    float        box_lo[3];
    float        box_scale[3];  // 1 / extent of the scene box
    uint32_t     guide_depth;   // vertices of depth below this sample the guide
    uint32_t     guide_medium;  // scattering vertices in the medium sample the guide too
    // hero-wavelength paths: four wavelengths each; throughput[0..2] hold
    // the first three, path_spec[path].x the fourth and .y the hero number
    uint32_t     spectral;
    float        spec_minv[9];  // the film's response: inverse Gram matrix of the bands
    float2*      path_spec;     // per path slot
    float*       nee_thr4;      // per queue entry: the vertex's fourth throughput
    float*       shadow_c4;     // per shadow entry: the fourth wavelength's contribution
    uint32_t     record_depth;  // vertices of depth below this record samples
    uint32_t     record_capacity;
    float*       train_x;       // 9 per record
    float*       train_dir;     // 3 per record
    float*       train_w;       // 1 per record
    float*       rec_t;         // throughput before the direction sample
    float*       rec_r;         // path radiance before the next vertex
    uint32_t*    rec_path;      // path of each record slot, Null: no record this wave
    uint32_t*    rec_block;     // per 256 record slots: valid records, then their first rank
};

__device__ __forceinline__ float offset_bits(float p, int32_t o) {
    return __int_as_float(__float_as_int(p) + (p < 0.f ? -o : o));
}

// offset_ray, core/scene/ray_offset.inl:18-29
__device__ __forceinline__ V3 offset_ray(V3 p, V3 n) {
    float const   origin      = 1.f / 32.f;
    float const   float_scale = 1.f / 65536.f;
    float const   int_scale   = 256.f;
    int32_t const ox          = int32_t(int_scale * n.x);
    int32_t const oy          = int32_t(int_scale * n.y);
    int32_t const oz          = int32_t(int_scale * n.z);
    return {fabsf(p.x) < origin ? __builtin_fmaf(float_scale, n.x, p.x) : offset_bits(p.x, ox),
            fabsf(p.y) < origin ? __builtin_fmaf(float_scale, n.y, p.y) : offset_bits(p.y, oy),
            fabsf(p.z) < origin ? __builtin_fmaf(float_scale, n.z, p.z) : offset_bits(p.z, oz)};
}

__device__ __forceinline__ V3 vertex(Wave const& w, uint32_t i) {
    float const* p = w.positions + 3 * size_t(i);
    return {p[0], p[1], p[2]};
}

// A closest hit of either scene kind: world triangles only, or instances.
struct Hit {
    float    t;
    float    u;
    float    v;
    uint32_t triangle;  // Null on a miss
    uint32_t instance;  // Null in a scene of world triangles
    uint32_t slot;      // the triangle's slot in its mesh tree (instances)
};

__device__ __forceinline__ Hit load_hit(Wave const& w, uint32_t i) {
    if (w.instanced) {
        mi_instance_hit const h = w.ihits[i];
        return {h.t, h.u, h.v, h.triangle, h.instance, h.slot};
    }
    mi_hit const h = w.hits[i];
    return {h.t, h.u, h.v, h.triangle, dev::Null, dev::Null};
}

// Point and geometric normal of a (non-missed) hit, its material and light.
struct Surface_point {
    V3       p;
    V3       n;   // geometric normal
    V3       ns;  // shading normal: the vertex normals interpolated, else n
    uint32_t material;
    uint32_t light;  // Null unless an emitter of the light tree
};

__device__ Surface_point surface_point(Wave const& w, Hit const& h) {
    Surface_point sp;
    if (dev::Null == h.instance || h.instance == w.emitter_instance) {
        // a world triangle (the instance of the world triangles is the identity)
        uint32_t const tri = h.triangle;
        V3 const       a   = vertex(w, w.indices[3 * size_t(tri) + 0]);
        V3 const       b   = vertex(w, w.indices[3 * size_t(tri) + 1]);
        V3 const       c   = vertex(w, w.indices[3 * size_t(tri) + 2]);
        float const    bw  = (1.f - h.u) - h.v;
        sp.p               = (bw * a + h.u * b) + h.v * c;
        sp.n               = dev::normalize(dev::cross(b - a, c - a));
        sp.ns              = sp.n;
        if (w.normals) {
            // triangle::interpolate_data (triangle_primitive_mt.inl:274-287):
            // (1 - u - v) n_a + u n_b + v n_c, normalized
            auto const nv = [&w](uint32_t k) {
                float const* q = w.normals + 3 * size_t(k);
                return v3(q[0], q[1], q[2]);
            };
            V3 const na = nv(w.indices[3 * size_t(tri) + 0]);
            V3 const nb = nv(w.indices[3 * size_t(tri) + 1]);
            V3 const nc = nv(w.indices[3 * size_t(tri) + 2]);
            sp.ns       = dev::normalize((bw * na + h.u * nb) + h.v * nc);
        }
        sp.material        = w.tri_material[tri];
        sp.light           = w.tri_light[tri];
        return sp;
    }
    // a mesh instance: the triangle interpolated in object space and moved to
    // world space, its normal normalized, transformed and normalized again
    // (Mesh::intersect, triangle_mesh.cpp:379-392; Indexed_data::normal,
    // triangle_bvh_indexed_data.inl:282-293)
    Mesh_geometry const g  = w.mesh_geo[w.tlas.instances[h.instance].mesh];
    auto const          at = [&g](uint32_t k) {
        float const* q = g.positions + 3 * size_t(k);
        return v3(q[0], q[1], q[2]);
    };
    V3 const      a  = at(g.indices[3 * size_t(h.triangle) + 0]);
    V3 const      b  = at(g.indices[3 * size_t(h.triangle) + 1]);
    V3 const      c  = at(g.indices[3 * size_t(h.triangle) + 2]);
    float const   bw = (1.f - h.u) - h.v;
    V3 const      po = (bw * a + h.u * b) + h.v * c;
    V3 const      no = dev::normalize(dev::cross(b - a, c - a));
    float const*  m  = w.inst_o2w + 16 * size_t(h.instance);
    dev::F3 const pw = dev::xform_point(m, {po.x, po.y, po.z});
    dev::F3 const nw = dev::xform_vector(m, {no.x, no.y, no.z});
    sp.p             = v3(pw.x, pw.y, pw.z);
    sp.n             = dev::normalize(v3(nw.x, nw.y, nw.z));
    sp.ns            = sp.n;
    if (g.normals) {
        // the mesh's vertex normals interpolated in object space, normalized,
        // then moved to world space as the geometric normal is
        auto const nv = [&g](uint32_t k) {
            float const* q = g.normals + 3 * size_t(k);
            return v3(q[0], q[1], q[2]);
        };
        V3 const      na = nv(g.indices[3 * size_t(h.triangle) + 0]);
        V3 const      nb = nv(g.indices[3 * size_t(h.triangle) + 1]);
        V3 const      nc = nv(g.indices[3 * size_t(h.triangle) + 2]);
        V3 const      so = dev::normalize((bw * na + h.u * nb) + h.v * nc);
        dev::F3 const sw = dev::xform_vector(m, {so.x, so.y, so.z});
        sp.ns            = dev::normalize(v3(sw.x, sw.y, sw.z));
    }
    sp.material      = w.inst_material[h.instance];
    sp.light         = dev::Null;
    return sp;
}

template <int NC>
__device__ __forceinline__ float max_channel(float const t[NC]) {
    float m = t[0];
    for (int c = 1; c < NC; ++c) {
        m = fmaxf(m, t[c]);
    }
    return m;
}

__device__ __forceinline__ float max_component(float const t[3]) {
    return fmaxf(t[0], fmaxf(t[1], t[2]));
}

// power_heuristic, integrator_helper.hpp:32-35
__device__ __forceinline__ float power_heuristic(float f, float g) {
    float const f2 = f * f;
    return f2 / (f2 + g * g);
}

// predivided_power_heuristic, integrator_helper.hpp:38-40
__device__ __forceinline__ float predivided_power_heuristic(float f, float g) {
    return f / (f * f + g * g);
}

// Reserves n consecutive entries per thread of the block in a queue with one
// atomic per block; the threads' ranges follow in thread order.  Every thread
// of the block calls it (blocks of up to 1024 threads).  Same-address atomics
// serialize, so one per block instead of one per wave keeps them off the
// critical path.
__device__ __forceinline__ uint32_t block_queue_range(uint32_t n, uint32_t* counter) {
    __shared__ uint32_t wave_total[16];
    __shared__ uint32_t block_base;
    uint32_t const lane = threadIdx.x & 63;
    uint32_t const wave = threadIdx.x >> 6;
    uint32_t       inc  = n;  // inclusive prefix sum over the wave's lanes
    for (uint32_t d = 1; d < 64; d <<= 1) {
        uint32_t const up = __shfl_up(inc, d, 64);
        inc += lane >= d ? up : 0u;
    }
    if (63 == lane) {
        wave_total[wave] = inc;
    }
    __syncthreads();
    if (0 == threadIdx.x) {
        uint32_t total = 0;
        for (uint32_t k = 0; k < (blockDim.x >> 6); ++k) {
            total += wave_total[k];
        }
        block_base = total ? atomicAdd(counter, total) : 0u;
    }
    __syncthreads();
    uint32_t base = block_base;
    for (uint32_t k = 0; k < wave; ++k) {
        base += wave_total[k];
    }
    __syncthreads();  // the next call reuses wave_total and block_base
    return base + inc - n;
}

// Scattering vertex in the medium: Henyey-Greenstein phase function
constexpr uint32_t Kind_phase = 0x100;

// The material of a vertex: Lambert or Substitute (both one-sided), or the
// medium's phase function (volumetric::Sample, volumetric_sample.cpp:15-60).
struct Surface {
    dev::Material_closure sub;  // sub.wo is the outgoing direction for every kind
    uint32_t              kind;
    float                 g;  // phase anisotropy

    __device__ void evaluate(V3 wi, float out[4]) const {
        if (Kind_phase == kind) {
            float const ph = dev::phase_hg(dev::dot(sub.wo, wi), g);
            out[0] = out[1] = out[2] = out[3] = ph;
            return;
        }
        if (MI_MATERIAL_LAMBERT == kind) {
            // lambert::Iso::reflection times n.wi, lambert.inl:15-26
            float const n_dot_wi = dev::clamp_dot(sub.layer.n, wi);
            for (int c = 0; c < 3; ++c) {
                out[c] = n_dot_wi * (dev::Pi_inv * sub.albedo[c]);
            }
            out[3] = n_dot_wi * dev::Pi_inv;
            return;
        }
        sub.evaluate(wi, out);
    }

    __device__ void sample(float xl, float x0, float x1, dev::Bxdf_sample& s) const {
        if (Kind_phase == kind) {
            float pdf;
            s.wi      = dev::sample_phase(sub.wo, x0, x1, g, pdf);
            s.refl[0] = s.refl[1] = s.refl[2] = pdf;
            s.pdf     = pdf;
            s.type    = dev::Type_diffuse;
            return;
        }
        if (MI_MATERIAL_LAMBERT == kind) {
            // lambert::Iso::reflect, lambert.inl:28-45
            V3 const    is       = dev::hemisphere_cosine(x0, x1);
            V3 const    wi       = dev::normalize(sub.layer.tangent_to_world(is));
            float const n_dot_wi = dev::clamp_dot(sub.layer.n, wi);
            for (int c = 0; c < 3; ++c) {
                s.refl[c] = n_dot_wi * (dev::Pi_inv * sub.albedo[c]);
            }
            s.wi   = wi;
            s.pdf  = n_dot_wi * dev::Pi_inv;
            s.type = dev::Type_diffuse;
            return;
        }
        sub.sample(xl, x0, x1, s);
    }
};

// The image dimension of sprout's Golden_ratio sampler for every pixel of
// the camera grid: the generator started as Worker::render starts it per
// pixel (seed 0, sequence = the grid pixel (y + r) W' + x + r, iteration 0;
// rendering_worker.cpp:97-104), then Golden_ratio::generate_2D
// (sampler_golden_ratio.cpp:51-61): two draws, math::golden_ratio over spp
// points (sample_distribution.inl:83-147) and rnd::biased_shuffle
// (shuffle.hpp:12-32).  Nothing before the first camera sample draws from
// the pixel's generator, so these are the pixel_uv of sprout's samples.
// table is sample-major (spp rows of num_pixels), so lanes of neighbouring
// pixels touch neighbouring words.
__global__ void __launch_bounds__(256) golden_ratio_samples(float2* table, uint32_t num_pixels,
                                                            uint32_t spp) {
    uint32_t const pixel = blockIdx.x * blockDim.x + threadIdx.x;
    if (pixel >= num_pixels) {
        return;
    }
    auto at = [=](uint32_t i) -> float2& { return table[size_t(i) * num_pixels + pixel]; };
    constexpr float Step = 0.618033988749894f;

    // the start pair float2(random_float(), random_float()): g++, which builds
    // the reference, evaluates the two arguments right to left
    Rng rng;
    rng.start(0, pixel);
    float const r1 = rng.uniform();
    float const r0 = rng.uniform();

    // the r0 sequence in .y, keeping the index of its smallest value
    float    x     = r0;
    float    least = x;
    uint32_t idx   = 0;
    for (uint32_t i = 0; i < spp; ++i) {
        at(i).y = x;
        if (x < least) {
            least = x;
            idx   = i;
        }
        x += Step;
        if (x >= 1.f) {
            x -= 1.f;
        }
    }
    // read in the Fibonacci order from the smallest into .x
    uint32_t f = 1, fp = 1, parity = 0;
    while (f + fp < spp) {
        uint32_t const t = f;
        f += fp;
        fp = t;
        ++parity;
    }
    uint32_t const inc = (parity & 1u) ? f : fp;
    uint32_t const dec = (parity & 1u) ? fp : f;
    at(0).x            = at(idx).y;
    for (uint32_t i = 1; i < spp; ++i) {
        if (idx < dec) {
            idx += inc;
            if (idx >= spp) {
                idx -= dec;
            }
        } else {
            idx -= dec;
        }
        at(i).x = at(idx).y;
    }
    // the r1 sequence in .y
    float y = r1;
    for (uint32_t i = 0; i < spp; ++i) {
        at(i).y = y;
        y += Step;
        if (y >= 1.f) {
            y -= 1.f;
        }
    }
    // biased_shuffle: from the back, partner = high word of random * (i + 1)
    for (uint32_t i = spp - 1; i > 0; --i) {
        uint64_t const o = (uint64_t(rng.next()) * (uint64_t(i) + 1)) >> 32;
        float2 const   t = at(i);
        at(i)            = at(uint32_t(o));
        at(uint32_t(o))  = t;
    }
}

__global__ void set_base(uint32_t* counters, uint32_t first_sample) {
    counters[C_base] = first_sample;
    for (uint32_t k = 0; k < 6; ++k) {
        counters[C_totals + k] = 0;
    }
}

__global__ void advance_base(uint32_t* counters, uint32_t samples) {
    counters[C_base] += samples;
}

// Starts sample k of the frame batch (plane k / num_pixels, pixel k %
// num_pixels of the camera grid) in path slot `slot`: the camera sample and
// the path state; returns the camera ray.
__device__ mi_ray start_path(Wave const& w, uint32_t k, uint32_t slot) {
    uint32_t const pixel  = k % w.num_pixels;  // of the camera grid
    uint32_t const plane  = k / w.num_pixels;
    uint32_t const sample = w.counters[C_base] + plane;
    int32_t const  x      = int32_t(pixel % w.grid_w) - w.border;
    int32_t const  y      = int32_t(pixel / w.grid_w) - w.border;

    Path path;
    path.rng.start(w.seed, uint64_t(sample) * w.num_pixels + pixel);
    float jx = path.rng.uniform();
    float jy = path.rng.uniform();
    if (w.pixel_samples) {
        // the pixel's Golden_ratio sequence (the two draws above keep the
        // path's generator where the random sampler leaves it)
        float2 const ps = w.pixel_samples[size_t(sample) * w.num_pixels + pixel];
        jx              = ps.x;
        jy              = ps.y;
    }
    path.jitter[0] = jx;
    path.jitter[1] = jy;

    // Perspective::generate_ray (camera_perspective.cpp:47-80) with sprout's
    // contraction into fused multiply-adds, normalize's Newton step on the
    // correctly rounded 1/sqrt, then object_to_world_vector
    // (composed_transformation.inl:84-93)
    float const c0 = float(x) + jx;
    float const c1 = float(y) + jy;
    float       dc[3];
    for (int k = 0; k < 3; ++k) {
        dc[k] = __builtin_fmaf(c1, w.d_y[k], __builtin_fmaf(c0, w.d_x[k], w.left_top[k]));
    }
    float const dd = __builtin_fmaf(dc[2], dc[2], __builtin_fmaf(dc[0], dc[0], dc[1] * dc[1]));
    float const ri = 1.f / sqrtf(dd);
    float const ns = (0.5f * ri) * (3.f - (dd * ri) * ri);
    float const nx = dc[0] * ns;
    float const ny = dc[1] * ns;
    float const nz = dc[2] * ns;
    float const* r0 = w.rotation;
    float const* r1 = w.rotation + 3;
    float const* r2 = w.rotation + 6;
    V3 const     d  = v3(__builtin_fmaf(nz, r2[0], __builtin_fmaf(ny, r1[0], nx * r0[0])),
                         __builtin_fmaf(nz, r2[1], __builtin_fmaf(ny, r1[1], nx * r0[1])),
                         __builtin_fmaf(nz, r2[2], __builtin_fmaf(nx, r0[2], ny * r1[2])));

    for (int c = 0; c < 3; ++c) {
        path.throughput[c] = 1.f;
        path.radiance[c]   = 0.f;
    }
    path.bsdf_pdf = 0.f;
    path.flags    = Flag_singular | Flag_primary;
    path.record   = dev::Null;
    path.prev_n[0] = path.prev_n[1] = path.prev_n[2] = 0.f;
    path.depth    = 0;
    path.sample   = k;
    w.paths[slot] = path;
    if (w.spectral) {
        w.path_spec[slot] = make_float2(
            1.f, spectral::hero_u(w.seed, uint64_t(sample) * w.num_pixels + pixel));
    }

    mi_ray r;
    r.origin[0]    = w.position[0];
    r.origin[1]    = w.position[1];
    r.origin[2]    = w.position[2];
    r.min_t        = 0.f;
    r.direction[0] = d.x;
    r.direction[1] = d.y;
    r.direction[2] = d.z;
    r.max_t        = Ray_max_t;
    return r;
}

// The first rays of a wave (or of a regenerating batch): sample p in slot p
// for the first min(num_paths, batch_samples) slots.
__global__ void __launch_bounds__(256) raygen(Wave w) {
    uint32_t const live = min(w.num_paths, w.batch_samples);
    uint32_t const p    = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < live) {
        w.rays[0][p]     = start_path(w, p, p);
        w.ray_path[0][p] = p;
    }
    if (0 == p) {
        w.counters[C_queue0]                = live;
        w.counters[C_queue0 + Counter_line] = 0;
        w.counters[C_shadow]                = 0;
        w.counters[C_done]                  = 0;
        w.counters[C_next]                  = live;  // the next sample to start
    }
}

// Clears the queue the next shade fills and the shadow queue.
__global__ void reset_queues(uint32_t* counters, uint32_t out_q) {
    counters[C_queue0 + Counter_line * out_q] = 0;
    counters[C_shadow]                        = 0;
    counters[C_done]                          = 0;
}

// Ray totals for the statistics: the queue just traced and the shadow queue.
__global__ void tally(uint32_t* counters, uint32_t in_q) {
    uint64_t* totals = reinterpret_cast<uint64_t*>(counters + C_totals);
    totals[0] += counters[C_queue0 + Counter_line * in_q];
    totals[1] += counters[C_shadow];
}

// The collision coefficients per unit density at a path's four wavelengths.
__device__ __forceinline__ dev::Channel_cc<4> wavelength_channels(dev::Medium_view const& m,
                                                                  float hero) {
    spectral::Wavelengths wl;
    wl.set(hero);
    dev::Channel_cc<4> cc;
    wl.upsample4(m.cc_a, cc.a);
    wl.upsample4(m.cc_s, cc.s);
    return cc;
}

// Tracking_multi::integrate for every queued ray: the part of the ray before
// its surface hit that lies in the medium box is tracked; the tracking weight
// multiplies the path throughput, and a scattering point replaces the hit.
// Spectral: the channels are the path's four wavelengths.
template <bool Spectral>
__device__ __forceinline__ void medium_track_queue(Wave const& w, uint32_t in_q) {
    constexpr int  NC     = Spectral ? 4 : 3;
    uint32_t const count  = w.counters[C_queue0 + Counter_line * in_q];
    uint32_t const stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        mi_ray const r     = w.rays[in_q][i];
        Hit const    h     = load_hit(w, i);
        V3 const     o     = v3(r.origin[0], r.origin[1], r.origin[2]);
        V3 const     d     = v3(r.direction[0], r.direction[1], r.direction[2]);
        uint32_t     event = Medium_none;
        float        t0, t1;
        if (dev::box_overlap(w.medium, o, d, r.min_t, h.t, t0, t1)) {
            uint32_t const p     = w.ray_path[in_q][i];
            Path           path  = w.paths[p];
            float          t_out = 0.f;
            float          tw[NC];
            int            e;
            if constexpr (Spectral) {
                float2 sp = w.path_spec[p];
                e = dev::track_segment<4>(w.medium, wavelength_channels(w.medium, sp.y), o, d, t0,
                                          t1, path.depth, 0 != (path.flags & Flag_nearest),
                                          path.rng, t_out, tw);
                sp.x *= tw[3];
                w.path_spec[p] = sp;
            } else {
                e = dev::track_segment(w.medium, o, d, t0, t1, path.depth,
                                       0 != (path.flags & Flag_nearest), path.rng, t_out, tw);
            }
            for (int c = 0; c < 3; ++c) {
                path.throughput[c] *= tw[c];
            }
            if (1 == e) {
                event                  = Medium_scatter;
                w.medium_p[3 * i + 0] = o.x + t_out * d.x;
                w.medium_p[3 * i + 1] = o.y + t_out * d.y;
                w.medium_p[3 * i + 2] = o.z + t_out * d.z;
            } else if (2 == e) {
                event = Medium_absorbed;
            }
            w.paths[p] = path;
        }
        w.medium_event[i] = event;
    }
}

__global__ void __launch_bounds__(256) medium_track(Wave w, uint32_t in_q) {
    medium_track_queue<false>(w, in_q);
}

__global__ void __launch_bounds__(256) medium_track_spectral(Wave w, uint32_t in_q) {
    medium_track_queue<true>(w, in_q);
}

// Network inputs of every queued hit: position in the scene box, geometric
// normal and outgoing direction, each mapped to [0,1].  A scattering point in
// the medium has no normal: its normal inputs are those of a zero vector.
__global__ void __launch_bounds__(256) guide_inputs(Wave w, uint32_t in_q) {
    uint32_t const count  = w.counters[C_queue0 + Counter_line * in_q];
    uint32_t const stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        Hit const h       = load_hit(w, i);
        float*    xo      = w.guide_x + 9 * size_t(i);
        bool const scatter = w.medium.present && Medium_scatter == w.medium_event[i];
        if (dev::Null == h.triangle && !scatter) {
            for (int k = 0; k < 9; ++k) {
                xo[k] = 0.5f;
            }
            continue;
        }
        mi_ray const r = w.rays[in_q][i];
        V3           n = v3(0.f, 0.f, 0.f);
        float        p3[3];
        if (scatter) {
            for (int k = 0; k < 3; ++k) {
                p3[k] = w.medium_p[3 * size_t(i) + k];
            }
        } else {
            Surface_point const sp = surface_point(w, h);
            n                      = sp.n;
            p3[0]                  = sp.p.x;
            p3[1]                  = sp.p.y;
            p3[2]                  = sp.p.z;
        }
        for (int k = 0; k < 3; ++k) {
            xo[k] = dev::clampf((p3[k] - w.box_lo[k]) * w.box_scale[k], 0.f, 1.f);
        }
        xo[3] = 0.5f * (n.x + 1.f);
        xo[4] = 0.5f * (n.y + 1.f);
        xo[5] = 0.5f * (n.z + 1.f);
        xo[6] = 0.5f * (1.f - r.direction[0]);
        xo[7] = 0.5f * (1.f - r.direction[1]);
        xo[8] = 0.5f * (1.f - r.direction[2]);
    }
}

// One bounce of Pathtracer_MIS::integrate (pathtracer_mis.cpp:95-279) for
// every queued ray: emission with its MIS weight, Russian roulette, the light
// picks of next-event estimation (their rays are made by the nee kernel), the
// direction sample (the BSDF's, or with Guided the one-sample mixture of the
// guide and the BSDF) and the next ray.  The path's fields are read and
// written where they are used, so few stay in registers.  Spectral: the
// throughput has the path's four wavelengths; rgb reflectance and emission
// are upsampled to them and contributions reach the radiance through the
// film's response.
template <bool Guided, bool Spectral>
__device__ __forceinline__ void shade_queue(Wave const& w, uint32_t in_q) {
    constexpr int  NC     = Spectral ? 4 : 3;
    uint32_t const count  = w.counters[C_queue0 + Counter_line * in_q];
    uint32_t const out_q  = in_q ^ 1;
    uint32_t const stride = gridDim.x * blockDim.x;

    // block-uniform trip count: every lane reaches the wave-wide queue appends
    for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += stride) {
        uint32_t const i = base + threadIdx.x;

        bool     spawn     = false;
        bool     live      = false;  // a shading vertex that continues the path
        bool     in_medium = false;  // a scattering vertex of the medium
        uint32_t p         = 0;
        float    thr[NC];
        for (int c = 0; c < NC; ++c) {
            thr[c] = 0.f;
        }
        float    hero      = 0.f;  // spectral: the path's hero number
        uint32_t flags     = 0;
        Rng      rng;
        rng.state = rng.inc = 0;
        uint32_t material   = 0;  // of a surface vertex
        V3       hit_p      = v3(0.f, 0.f, 0.f);
        V3       geo_n      = v3(0.f, 0.f, 1.f);
        V3       shading_n  = geo_n;
        V3       wo         = v3(0.f, 0.f, 1.f);
        V3       p_o        = v3(0.f, 0.f, 0.f);
        // next-event estimation: the lights the tree picks at p_o
        // (sample_lights); counted here, picked once their queue slots are known
        uint32_t num_picks = 0;
        float    select    = 0.f;
        uint32_t prev_record = dev::Null;  // training record of the previous vertex's sample
        uint32_t depth       = 0;          // of this vertex
        bool     split       = false;

        if (i < count) {
            p                = w.ray_path[in_q][i];
            Path const* pp   = &w.paths[p];
            depth            = pp->depth;
            // sample_lights splits the light tree at the first bounces
            // (Pathtracer_MIS::splitting, pathtracer_mis.cpp:423-426)
            split            = w.split_lights && depth < Num_split_bounces;
            thr[0]           = pp->throughput[0];
            thr[1]           = pp->throughput[1];
            thr[2]           = pp->throughput[2];
            if constexpr (Spectral) {
                float2 const sp = w.path_spec[p];
                thr[3]          = sp.x;
                hero            = sp.y;
            }
            flags            = pp->flags;
            rng              = pp->rng;
            Hit const h      = load_hit(w, i);
            mi_ray const r   = w.rays[in_q][i];
            wo               = v3(-r.direction[0], -r.direction[1], -r.direction[2]);

            if (Guided) {
                prev_record = pp->record;
                if (dev::Null != prev_record) {
                    // radiance before this vertex: what follows is the previous sample's
                    float* rr = w.rec_r + 3 * size_t(prev_record);
                    rr[0]     = pp->radiance[0];
                    rr[1]     = pp->radiance[1];
                    rr[2]     = pp->radiance[2];
                }
            }

            uint32_t const event = w.medium.present ? w.medium_event[i] : Medium_none;
            bool           roulette = false;

            if (Medium_scatter == event) {
                // Event::Scatter: the vertex is the scattering point
                // (pathtracer_mis.cpp:215-250); no emission, the same depth
                // and roulette tests as a surface
                hit_p = v3(w.medium_p[3 * size_t(i) + 0], w.medium_p[3 * size_t(i) + 1],
                           w.medium_p[3 * size_t(i) + 2]);
                in_medium = true;
                live      = true;
                roulette  = 0 != depth;
            } else if (Medium_absorbed == event) {
                // Event::Abort: the tracking weight fell below Abort_epsilon
            } else if (dev::Null != h.triangle) {
                Surface_point const sp = surface_point(w, h);
                hit_p                  = sp.p;
                geo_n                  = sp.n;
                shading_n              = sp.ns;
                material               = sp.material;
                uint32_t const kind    = w.materials[sp.material].kind;
                live                   = true;

                if (MI_MATERIAL_LIGHT == kind) {
                    // first vertex: mat_sample.radiance(); later: connect_light
                    float const cos_l = dev::dot(geo_n, wo);
                    if (cos_l > 0.f) {
                        float          weight = 1.f;
                        uint32_t const l      = sp.light;
                        // an emitter light sampling cannot pick (zero power)
                        // is found by BSDF sampling alone: weight 1
                        if (0 != depth && !(flags & Flag_singular) && dev::Null != l) {
                            // Mesh::pdf times Tree::pdf at the previous vertex
                            V3 const prev_p = v3(r.origin[0], r.origin[1], r.origin[2]);
                            V3 const prev_n = v3(pp->prev_n[0], pp->prev_n[1], pp->prev_n[2]);
                            bool const translucent = 0 != (flags & Flag_translucent);
                            float const tree_pdf   = dev::light_pdf(w.tree, prev_p, prev_n, translucent, split, l);
                            // the triangle's pdf in its part's Primitive_tree
                            // times its relative area (Part::pdf, triangle_mesh.cpp:264-270)
                            uint32_t const    local = w.tri_local[h.triangle];
                            float const       ppdf  = dev::part_pdf(w.parts, l, prev_p, prev_n, translucent, local);
                            Light_info const& L = w.tree.lights[l];
                            Light_info const& T = w.parts.lights[w.parts.parts[l].light_offset + local];
                            float const ls_pdf  = (h.t * h.t) / (cos_l * L.area) * (ppdf * (L.area / T.area));
                            weight              = power_heuristic(pp->bsdf_pdf, ls_pdf * tree_pdf);
                        }
                        float const* e  = w.materials[sp.material].emission;
                        float*       rad = w.paths[p].radiance;
                        if constexpr (Spectral) {
                            spectral::Wavelengths wl;
                            wl.set(hero);
                            float el[4], c4[4], rgb[3];
                            wl.upsample4(e, el);
                            for (int k = 0; k < 4; ++k) {
                                c4[k] = thr[k] * (weight * el[k]);
                            }
                            wl.to_rgb(w.spec_minv, c4, rgb);
                            for (int k = 0; k < 3; ++k) {
                                rad[k] = rad[k] + rgb[k];
                            }
                            if (Guided && dev::Null != prev_record) {
                                float* rr = w.rec_r + 3 * size_t(prev_record);
                                for (int k = 0; k < 4; ++k) {
                                    c4[k] = thr[k] * ((1.f - weight) * el[k]);
                                }
                                wl.to_rgb(w.spec_minv, c4, rgb);
                                for (int k = 0; k < 3; ++k) {
                                    rr[k] -= rgb[k];
                                }
                            }
                        } else {
                            for (int k = 0; k < 3; ++k) {
                                rad[k] = rad[k] + thr[k] * (weight * e[k]);
                            }
                            if (Guided && dev::Null != prev_record) {
                                // the guide learns the whole incident radiance: the
                                // emission found by the sampled direction counts
                                // without its MIS weight
                                float* rr = w.rec_r + 3 * size_t(prev_record);
                                for (int k = 0; k < 3; ++k) {
                                    rr[k] -= thr[k] * ((1.f - weight) * e[k]);
                                }
                            }
                        }
                    }
                    live = false;
                } else {
                    roulette = 0 != depth;
                }
            }

            if (roulette) {
                if (depth >= w.max_bounces) {
                    live = false;
                } else if (depth >= w.min_bounces) {
                    // russian_roulette, integrator_helper.hpp:42-52: the
                    // continuation probability is the largest channel, and
                    // the throughput is divided by it even above one, which
                    // loses energy; sprout's throughput stays at or below one
                    // almost always.  A guided path's does not (the mixture
                    // pdf can be far below the BSDF's), so there the
                    // probability is capped at one, which keeps it unbiased.
                    float const q = Guided ? fminf(max_channel<NC>(thr), 1.f) : max_channel<NC>(thr);
                    if (rng.uniform() > q) {
                        live = false;
                    } else {
                        for (int k = 0; k < NC; ++k) {
                            thr[k] /= q;
                        }
                    }
                }
            }
            if (live && !in_medium && !(dev::dot(geo_n, wo) > 0.f)) {
                live = false;  // back face: evaluate and sample give nothing
            }
            if (live) {
                // offset_p: a surface vertex moves off the surface; a
                // scattering vertex uses the point itself and its tree
                // weighs lights over the whole sphere
                p_o = in_medium ? hit_p : offset_ray(hit_p, geo_n);
                if (w.tree.num_nodes) {
                    select    = rng.uniform();
                    num_picks = dev::count_lights(w.tree, p_o, split);
                }
                bool const linear = depth <= 1 || 0 != (flags & Flag_primary);
                flags             = linear ? (flags & ~Flag_nearest) : (flags | Flag_nearest);
            }
        }

        bool const guided_vertex = Guided && live && (!in_medium || 0 != w.guide_medium) &&
                                   depth < w.guide_depth;

        // next-event estimation (sample_lights / evaluate_light): a vertex's
        // shadow rays take consecutive queue slots, reserved with one atomic
        // per block.  The tree's picks go to those slots and the nee kernel
        // makes their rays from the random numbers that follow the selection
        // draw; the path's generator skips them here.
        uint32_t const sbase = block_queue_range(num_picks, &w.counters[C_shadow]);
        if (i < count) {
            w.shadow_range[i] = make_uint2(sbase, num_picks);
        }
        if (num_picks) {
            Nee_vertex v;
            v.rng_state = rng.state;
            v.rng_inc   = rng.inc;
            v.p[0]      = p_o.x;
            v.p[1]      = p_o.y;
            v.p[2]      = p_o.z;
            v.path      = p;
            v.n[0]      = geo_n.x;
            v.n[1]      = geo_n.y;
            v.n[2]      = geo_n.z;
            v.material  = in_medium ? Kind_phase : material;
            v.ns[0]     = shading_n.x;
            v.ns[1]     = shading_n.y;
            v.ns[2]     = shading_n.z;
            v.wo[0]     = wo.x;
            v.wo[1]     = wo.y;
            v.wo[2]     = wo.z;
            v.flags     = (guided_vertex ? Nee_guided : 0u) | (in_medium ? Nee_medium : 0u);
            for (int c = 0; c < 3; ++c) {
                v.throughput[c] = thr[c];
            }
            if constexpr (Spectral) {
                w.nee_thr4[i] = thr[3];
            }
            v.depth  = depth;
            w.nee[i] = v;
            uint32_t const made = dev::for_each_light(
                w.tree, p_o, geo_n, in_medium, select, split, [&](uint32_t k, dev::Light_pick const pick) {
                    if (k < num_picks) {
                        w.picks[sbase + k] = Pick_slot{pick.light, pick.pdf, i, k};
                    }
                });
            for (uint32_t k = made; k < num_picks; ++k) {
                // unreachable while both descents agree: a slot without a ray
                w.picks[sbase + k] = Pick_slot{dev::Null, 0.f, i, k};
            }
            rng.advance(3 * uint64_t(num_picks));  // rp, u1, u2 per pick
        }

        mi_ray next_ray;
        bool   record   = false;
        float  rec_t[3] = {0.f, 0.f, 0.f};
        if (live) {
            Surface s;
            if (in_medium) {
                s.kind   = Kind_phase;
                s.g      = w.medium.g;
                s.sub.wo = wo;
            } else {
                Render_material const m = w.materials[material];
                s.kind                  = m.kind;
                s.sub.set(dev::frame_from_normal(shading_n), geo_n, wo, m.color, m.roughness, m.metallic,
                          m.ior, 1.f, w.e_table);
            }
            // BSDF importance sample; guided: one-sample mixture of the guide
            // (probability fraction) and the BSDF
            dev::Bxdf_sample bs;
            if (guided_vertex) {
                // the mixture goes to memory: this sample and the MIS weights
                // of the vertex's shadow rays read it there
                float* const mix = w.guide_mix + dev::Vmf_mixture::Stored * size_t(i);
                {
                    dev::Vmf_mixture m;
                    m.set(w.guide_raw + 32 * size_t(i));
                    m.store(mix);
                }
                float const pick = rng.uniform();
                float const xl   = rng.uniform();
                float const x0   = rng.uniform();
                float const x1   = rng.uniform();
                if (pick < w.fraction) {
                    V3 const gw = dev::vmf_mixture_sample(mix, xl, x0, x1);
                    float    f[4];
                    s.evaluate(gw, f);
                    bs.wi      = gw;
                    bs.refl[0] = f[0];
                    bs.refl[1] = f[1];
                    bs.refl[2] = f[2];
                    bs.pdf     = f[3];
                    bs.type    = dev::Type_diffuse;
                    if (!in_medium && !(dev::dot(geo_n, gw) > 0.f)) {
                        bs.pdf = 0.f;  // under the surface: the reflection is zero
                    }
                } else {
                    s.sample(xl, x0, x1, bs);
                }
                if (bs.pdf > 0.f) {
                    bs.pdf = w.fraction * dev::vmf_mixture_pdf(mix, bs.wi) + (1.f - w.fraction) * bs.pdf;
                }
                if (bs.pdf > 0.f && depth < w.record_depth) {
                    record = true;
                    if constexpr (Spectral) {
                        // the training weight's throughput: the wavelengths' mean
                        float const mean = (((thr[0] + thr[1]) + thr[2]) + thr[3]) * 0.25f;
                        rec_t[0] = rec_t[1] = rec_t[2] = mean;
                    } else {
                        rec_t[0] = thr[0];
                        rec_t[1] = thr[1];
                        rec_t[2] = thr[2];
                    }
                }
            } else {
                float const xl = rng.uniform();
                float const x0 = rng.uniform();
                float const x1 = rng.uniform();
                s.sample(xl, x0, x1, bs);
            }
            if (bs.pdf > 0.f) {
                if (dev::Type_specular == bs.type) {
                    flags |= Flag_singular;
                } else {
                    flags &= ~(Flag_singular | Flag_primary);
                }
                if (in_medium) {
                    flags |= Flag_translucent;
                } else {
                    flags &= ~Flag_translucent;
                }
                if constexpr (Spectral) {
                    spectral::Wavelengths wl;
                    wl.set(hero);
                    float rl[4];
                    wl.upsample4(bs.refl, rl);
                    for (int k = 0; k < 4; ++k) {
                        thr[k] *= rl[k] / bs.pdf;
                    }
                } else {
                    for (int k = 0; k < 3; ++k) {
                        thr[k] *= bs.refl[k] / bs.pdf;
                    }
                }
                Path* const pp = &w.paths[p];
                pp->bsdf_pdf   = bs.pdf;
                pp->depth      = depth + 1;
                pp->prev_n[0]  = geo_n.x;
                pp->prev_n[1]  = geo_n.y;
                pp->prev_n[2]  = geo_n.z;
                // a scattering vertex spawns from the point itself (subsurface offset_p)
                V3 const o = in_medium ? hit_p
                                       : offset_ray(hit_p, dev::dot(geo_n, bs.wi) > 0.f ? geo_n : -geo_n);
                next_ray.origin[0]    = o.x;
                next_ray.origin[1]    = o.y;
                next_ray.origin[2]    = o.z;
                next_ray.min_t        = 0.f;
                next_ray.direction[0] = bs.wi.x;
                next_ray.direction[1] = bs.wi.y;
                next_ray.direction[2] = bs.wi.z;
                next_ray.max_t        = Ray_max_t;
                spawn                 = true;
            }
        }

        if (i < count) {
            Path* const pp    = &w.paths[p];
            pp->throughput[0] = thr[0];
            pp->throughput[1] = thr[1];
            pp->throughput[2] = thr[2];
            if constexpr (Spectral) {
                w.path_spec[p].x = thr[3];
            }
            pp->flags         = flags;
            pp->rng           = rng;
            if (Guided) {
                // a fixed slot per (path, depth): the records, and so the
                // training, do not depend on the order the queue was filled in
                uint32_t       rec    = dev::Null;
                uint64_t const slot64 = uint64_t(p) * w.record_depth + depth;
                if (record && spawn && slot64 < w.record_capacity) {
                    uint32_t const slot = uint32_t(slot64);
                    float const*   xi   = w.guide_x + 9 * size_t(i);
                    float*         xo   = w.train_x + 9 * size_t(slot);
                    for (int k = 0; k < 9; ++k) {
                        xo[k] = xi[k];
                    }
                    w.train_dir[3 * size_t(slot) + 0] = next_ray.direction[0];
                    w.train_dir[3 * size_t(slot) + 1] = next_ray.direction[1];
                    w.train_dir[3 * size_t(slot) + 2] = next_ray.direction[2];
                    for (int k = 0; k < 3; ++k) {
                        w.rec_t[3 * size_t(slot) + k] = rec_t[k];
                        w.rec_r[3 * size_t(slot) + k] = pp->radiance[k];
                    }
                    w.rec_path[slot] = p;
                    rec              = slot;
                }
                pp->record = rec;
            }
        }

        uint32_t const ray_slot = block_queue_range(spawn ? 1u : 0u, &w.counters[C_queue0 + Counter_line * out_q]);
        if (spawn) {
            w.rays[out_q][ray_slot]     = next_ray;
            w.ray_path[out_q][ray_slot] = p;
        }
        if (w.regenerate) {
            // the path ends here: its slot goes to the next sample once the
            // vertex's shadow rays are resolved
            bool const     ended = i < count && !spawn;
            uint32_t const d     = block_queue_range(ended ? 1u : 0u, &w.counters[C_done]);
            if (ended) {
                w.done[d] = p;
            }
        }
    }
}

// 127 VGPRs: four waves per SIMD
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) shade(Wave w,
                                                                                   uint32_t in_q) {
    shade_queue<false, false>(w, in_q);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
shade_spectral(Wave w, uint32_t in_q) {
    shade_queue<false, true>(w, in_q);
}

__global__ void __launch_bounds__(256) shade_guided_spectral(Wave w, uint32_t in_q) {
    shade_queue<true, true>(w, in_q);
}

__global__ void __launch_bounds__(256) shade_guided(Wave w, uint32_t in_q) {
    shade_queue<true, false>(w, in_q);
}

// Regenerating waves: every path that ended this bounce (its last vertex's
// shadow rays now resolved) leaves its radiance in the batch's sample buffer,
// and its slot starts the batch's next sample, whose camera ray joins the
// queue the next bounce traces.  One atomic per block for the samples and
// one for the queue.
__global__ void __launch_bounds__(256) regenerate(Wave w, uint32_t out_q) {
    uint32_t const count  = w.counters[C_done];
    uint32_t const stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += stride) {  // block-uniform
        uint32_t const i    = base + threadIdx.x;
        bool const     have = i < count;
        uint32_t       slot = 0;
        if (have) {
            slot                = w.done[i];
            Path const&    path = w.paths[slot];
            float* const   out  = w.sample_rad + 3 * size_t(path.sample);
            out[0]              = path.radiance[0];
            out[1]              = path.radiance[1];
            out[2]              = path.radiance[2];
        }
        uint32_t const k     = block_queue_range(have ? 1u : 0u, &w.counters[C_next]);
        bool const     start = have && k < w.batch_samples;
        mi_ray         ray;
        if (start) {
            ray = start_path(w, k, slot);
        }
        uint32_t const q = block_queue_range(start ? 1u : 0u, &w.counters[C_queue0 + Counter_line * out_q]);
        if (start) {
            w.rays[out_q][q]     = ray;
            w.ray_path[out_q][q] = slot;
        }
    }
}

// Regenerating waves: the batch's samples summed per pixel in the order the
// waves of spw planes add them (the same FP64 sums as resolve).
__global__ void __launch_bounds__(256) resolve_samples(Wave w, uint32_t planes, uint32_t spw,
                                                       double* accum) {
    uint32_t const pixel = blockIdx.x * blockDim.x + threadIdx.x;
    if (pixel >= w.num_pixels) {
        return;
    }
    double a[3] = {accum[3 * size_t(pixel) + 0], accum[3 * size_t(pixel) + 1],
                   accum[3 * size_t(pixel) + 2]};
    for (uint32_t k0 = 0; k0 < planes; k0 += spw) {
        double s[3] = {0.0, 0.0, 0.0};
        for (uint32_t k = k0; k < k0 + spw; ++k) {
            float const* r = w.sample_rad + 3 * (size_t(k) * w.num_pixels + pixel);
            for (int c = 0; c < 3; ++c) {
                s[c] += double(r[c]);
            }
        }
        for (int c = 0; c < 3; ++c) {
            a[c] += s[c];
        }
    }
    for (int c = 0; c < 3; ++c) {
        accum[3 * size_t(pixel) + c] = a[c];
    }
}

// The shadow ray of every light pick of the bounce (evaluate_light,
// pathtracer_mis.cpp:323-358), one thread per pick: Mesh::sample picks a
// triangle of the part and a point on it (3 random numbers of the path's
// generator, pick k taking numbers 3k..3k+2 after the vertex's selection
// draw), the vertex's reflection toward it and the MIS weight against the
// vertex's direction sampling (with the guide's mixture at a guided vertex).
// A ray with nothing to add gets an empty interval.
// Spectral: the contribution per wavelength (the fourth in shadow_c4), the
// light's emission and the reflection upsampled to the path's wavelengths.
template <bool Spectral>
__device__ __forceinline__ void nee_queue(Wave const& w) {
    uint32_t const count  = w.counters[C_shadow];
    uint32_t const stride = gridDim.x * blockDim.x;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < count; j += stride) {
        Pick_slot const  pick = w.picks[j];
        Nee_vertex const v    = w.nee[pick.vertex];
        bool             want = false;
        mi_ray           shadow_ray;
        Shadow           sh;
        float            c4 = 0.f;
        if (dev::Null != pick.light) {
            Rng rng;
            rng.state = v.rng_state;
            rng.inc   = v.rng_inc;
            rng.advance(3 * uint64_t(pick.k));
            bool const in_medium = 0 != (v.flags & Nee_medium);
            V3 const   p_o       = v3(v.p[0], v.p[1], v.p[2]);
            V3 const   geo_n     = v3(v.n[0], v.n[1], v.n[2]);
            Surface    s;
            if (in_medium) {
                s.kind   = Kind_phase;
                s.g      = w.medium.g;
                s.sub.wo = v3(v.wo[0], v.wo[1], v.wo[2]);
            } else {
                Render_material const m = w.materials[v.material];
                s.kind                  = m.kind;
                s.sub.set(dev::frame_from_normal(v3(v.ns[0], v.ns[1], v.ns[2])), geo_n,
                          v3(v.wo[0], v.wo[1], v.wo[2]),
                          m.color, m.roughness, m.metallic, m.ior, 1.f, w.e_table);
            }

            Light_info const L = w.tree.lights[pick.light];  // the part
            // Mesh::sample (triangle_mesh.cpp:462-516): the part's tree
            // picks a triangle, then a point on it
            float const           rp = rng.uniform();
            dev::Light_pick const tp = dev::part_random(w.parts, pick.light, p_o, geo_n, in_medium, rp);
            Light_info const      T  = w.parts.lights[w.parts.parts[pick.light].light_offset + tp.light];
            uint32_t const        lt = T.triangle;
            float const           u1 = rng.uniform();
            float const           u2 = rng.uniform();

            V3 const la = vertex(w, w.indices[3 * size_t(lt) + 0]);
            V3 const lb = vertex(w, w.indices[3 * size_t(lt) + 1]);
            V3 const lc = vertex(w, w.indices[3 * size_t(lt) + 2]);
            // sample_triangle_uniform, base/math/sampling.inl:39-43
            float const su = sqrtf(u1);
            float const lu = 1.f - su;
            float const lv = u2 * su;
            float const lw = (1.f - lu) - lv;
            V3 const    lp = (lw * la + lu * lb) + lv * lc;
            V3 const    ln = dev::normalize(dev::cross(lb - la, lc - la));

            V3 const    axis = offset_ray(lp, ln) - p_o;
            float const sl   = dev::dot(axis, axis);
            float const dist = sqrtf(sl);
            V3 const    wi   = v3(axis.x / dist, axis.y / dist, axis.z / dist);
            float const cl   = -dev::dot(ln, wi);
            if (cl >= dev::Dot_min) {
                float f[4];
                s.evaluate(wi, f);
                float const light_pdf = (sl / (cl * L.area)) * (tp.pdf * (L.area / T.area)) * pick.pdf;
                float       dir_pdf   = f[3];
                if (v.flags & Nee_guided) {
                    float const* mix = w.guide_mix + dev::Vmf_mixture::Stored * size_t(pick.vertex);
                    dir_pdf = w.fraction * dev::vmf_mixture_pdf(mix, wi) + (1.f - w.fraction) * f[3];
                }
                float const           weight = predivided_power_heuristic(light_pdf, dir_pdf);
                Render_material const lm     = w.materials[w.tri_material[lt]];
                if constexpr (Spectral) {
                    spectral::Wavelengths wl;
                    wl.set(w.path_spec[v.path].y);
                    float el[4], fl[4];
                    wl.upsample4(lm.emission, el);
                    wl.upsample4(f, fl);
                    float const t4[4] = {v.throughput[0], v.throughput[1], v.throughput[2],
                                         w.nee_thr4[pick.vertex]};
                    for (int c = 0; c < 3; ++c) {
                        sh.contrib[c] = t4[c] * (weight * (el[c] * fl[c]));
                    }
                    c4 = t4[3] * (weight * (el[3] * fl[3]));
                } else {
                    for (int c = 0; c < 3; ++c) {
                        sh.contrib[c] = v.throughput[c] * (weight * (lm.emission[c] * f[c]));
                    }
                }
                sh.path                 = v.path;
                shadow_ray.origin[0]    = p_o.x;
                shadow_ray.origin[1]    = p_o.y;
                shadow_ray.origin[2]    = p_o.z;
                shadow_ray.min_t        = 0.f;
                shadow_ray.direction[0] = wi.x;
                shadow_ray.direction[1] = wi.y;
                shadow_ray.direction[2] = wi.z;
                shadow_ray.max_t        = dist;
                want = sh.contrib[0] > 0.f || sh.contrib[1] > 0.f || sh.contrib[2] > 0.f ||
                       c4 > 0.f;
            }
        }
        if (!want) {
            // nothing to add: an empty interval, resolved as not occluded
            shadow_ray.origin[0] = shadow_ray.origin[1] = shadow_ray.origin[2] = 0.f;
            shadow_ray.direction[0] = 0.f;
            shadow_ray.direction[1] = 0.f;
            shadow_ray.direction[2] = 1.f;
            shadow_ray.min_t        = 1.f;
            shadow_ray.max_t        = 0.f;
            sh.contrib[0] = sh.contrib[1] = sh.contrib[2] = 0.f;
            sh.path                 = v.path;
            c4                      = 0.f;
        }
        w.shadow_rays[j] = shadow_ray;
        w.shadows[j]     = sh;
        if constexpr (Spectral) {
            w.shadow_c4[j] = c4;
        }
    }
}

__global__ void __launch_bounds__(256) nee(Wave w) {
    nee_queue<false>(w);
}

__global__ void __launch_bounds__(256) nee_spectral(Wave w) {
    nee_queue<true>(w);
}

// Adds the unoccluded light contributions of every vertex shaded this
// bounce to its path, in the order the lights were picked.
// With a medium, the part of each unoccluded shadow ray inside its box is
// attenuated by ratio tracking (Worker::transmitted: visibility, then
// transmittance, rendering_worker.cpp:143-153), drawing from the path's
// generator after the vertex's direction sample.
// Spectral: transmittance per wavelength, then the film's response.
template <bool Spectral>
__device__ __forceinline__ void shadow_resolve_queue(Wave const& w, uint32_t in_q) {
    constexpr int  NC     = Spectral ? 4 : 3;
    uint32_t const count  = w.counters[C_queue0 + Counter_line * in_q];
    uint32_t const stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        uint2 const range = w.shadow_range[i];
        if (0 == range.y) {
            continue;
        }
        uint32_t const path        = w.ray_path[in_q][i];
        Path* const    p           = &w.paths[path];
        Rng            rng         = p->rng;
        float          radiance[3] = {p->radiance[0], p->radiance[1], p->radiance[2]};
        spectral::Wavelengths wl;
        dev::Channel_cc<NC>   cc;
        if constexpr (Spectral) {
            wl.set(w.path_spec[path].y);
            wl.upsample4(w.medium.cc_a, cc.a);
            wl.upsample4(w.medium.cc_s, cc.s);
        } else {
            cc = dev::rgb_channels(w.medium);
        }
        for (uint32_t k = 0; k < range.y; ++k) {
            if (w.occluded[range.x + k]) {
                continue;
            }
            Shadow const s = w.shadows[range.x + k];
            float        tr[NC];
            for (int c = 0; c < NC; ++c) {
                tr[c] = 1.f;
            }
            if (w.medium.present) {
                mi_ray const r = w.shadow_rays[range.x + k];
                V3 const     o = v3(r.origin[0], r.origin[1], r.origin[2]);
                V3 const     d = v3(r.direction[0], r.direction[1], r.direction[2]);
                float        t0, t1;
                if (r.min_t < r.max_t && dev::box_overlap(w.medium, o, d, r.min_t, r.max_t, t0, t1)) {
                    if (!dev::transmittance<NC>(w.medium, cc, o, d, t0, t1, w.nee[i].depth,
                                                0 != (p->flags & Flag_nearest), rng, tr)) {
                        continue;
                    }
                }
            }
            if constexpr (Spectral) {
                float const c4[4] = {tr[0] * s.contrib[0], tr[1] * s.contrib[1],
                                     tr[2] * s.contrib[2], tr[3] * w.shadow_c4[range.x + k]};
                float       rgb[3];
                wl.to_rgb(w.spec_minv, c4, rgb);
                for (int c = 0; c < 3; ++c) {
                    radiance[c] += rgb[c];
                }
            } else {
                for (int c = 0; c < 3; ++c) {
                    radiance[c] += tr[c] * s.contrib[c];
                }
            }
        }
        for (int c = 0; c < 3; ++c) {
            p->radiance[c] = radiance[c];
        }
        p->rng = rng;
    }
}

__global__ void __launch_bounds__(256) shadow_resolve(Wave w, uint32_t in_q) {
    shadow_resolve_queue<false>(w, in_q);
}

__global__ void __launch_bounds__(256) shadow_resolve_spectral(Wave w, uint32_t in_q) {
    shadow_resolve_queue<true>(w, in_q);
}

// Marks every record slot of the wave empty; the shading kernels fill the
// slot path * record_depth + depth of each vertex that records.
__global__ void __launch_bounds__(256) clear_records(Wave w) {
    uint32_t const stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < w.record_capacity; i += stride) {
        w.rec_path[i] = dev::Null;
    }
}

// Training weight of every record: average over rgb of (radiance the path
// gathered after the record's vertex, the emitter its direction hit counted
// without MIS weight) / (throughput before its direction sample) = incident
// radiance x reflection / sampling pdf.
__global__ void __launch_bounds__(256) finish_records(Wave w) {
    uint32_t const stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < w.record_capacity; i += stride) {
        uint32_t const path = w.rec_path[i];
        if (dev::Null == path) {
            continue;
        }
        Path const& p   = w.paths[path];
        float       sum = 0.f;
        for (int k = 0; k < 3; ++k) {
            float const t = w.rec_t[3 * size_t(i) + k];
            float const d = p.radiance[k] - w.rec_r[3 * size_t(i) + k];
            sum += t > 0.f ? fmaxf(d, 0.f) / t : 0.f;
        }
        float const v = sum * (1.f / 3.f);
        w.train_w[i]  = isfinite(v) ? v : 0.f;
    }
}

// Rank of a valid record slot among the valid slots of its block of 256
// (slot order), and the block's count.  Every thread of the block calls it.
__device__ __forceinline__ uint32_t block_rank(bool valid, uint32_t* block_total) {
    __shared__ uint32_t wave_counts[4];
    uint32_t const lane  = threadIdx.x & 63;
    uint32_t const wave  = threadIdx.x >> 6;
    uint64_t const mask  = __ballot(valid);
    if (0 == lane) {
        wave_counts[wave] = uint32_t(__popcll(mask));
    }
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t k = 0; k < wave; ++k) {
        before += wave_counts[k];
    }
    *block_total = (wave_counts[0] + wave_counts[1]) + (wave_counts[2] + wave_counts[3]);
    return before + uint32_t(__popcll(mask & ((uint64_t(1) << lane) - 1)));
}

// Valid records per block of 256 slots (blocks of 256 threads, one slot each).
__global__ void __launch_bounds__(256) count_records(Wave w) {
    uint32_t const i     = blockIdx.x * 256 + threadIdx.x;
    bool const     valid = i < w.record_capacity && dev::Null != w.rec_path[i];
    uint32_t       total = 0;
    block_rank(valid, &total);
    if (0 == threadIdx.x) {
        w.rec_block[blockIdx.x] = total;
    }
}

// Exclusive prefix sums of the block counts, in block order (one block of
// 1024 threads, each over a contiguous run); the total is the wave's record
// count.
__global__ void __launch_bounds__(1024) scan_record_blocks(Wave w, uint32_t num_blocks) {
    __shared__ uint32_t sums[1024];
    uint32_t const t     = threadIdx.x;
    uint32_t const run   = (num_blocks + 1023) / 1024;
    uint32_t const begin = min(t * run, num_blocks);
    uint32_t const end   = min(begin + run, num_blocks);
    uint32_t       local = 0;
    for (uint32_t b = begin; b < end; ++b) {
        local += w.rec_block[b];
    }
    sums[t] = local;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        uint32_t const add = t >= d ? sums[t - d] : 0u;
        __syncthreads();
        sums[t] += add;
        __syncthreads();
    }
    uint32_t offset = sums[t] - local;
    for (uint32_t b = begin; b < end; ++b) {
        uint32_t const c = w.rec_block[b];
        w.rec_block[b]   = offset;
        offset += c;
    }
    if (1023 == t) {
        w.counters[C_records] = sums[1023];
    }
}

// Moves the valid records, in slot order, to the shuffled copy: record j
// (its rank among the valid slots) goes to (j P + salt) mod count,
// P = 2^31 - 1 (prime, so a bijection for every count below it).  The slots
// follow the pixels; consecutive training slices of the shuffled copy mix
// the whole image.  Blocks of 256 threads, one slot each, as count_records.
__global__ void __launch_bounds__(256) shuffle_records(Wave w, uint32_t const* step, float* sx,
                                                       float* sd, float* sw) {
    uint32_t const i     = blockIdx.x * 256 + threadIdx.x;
    bool const     valid = i < w.record_capacity && dev::Null != w.rec_path[i];
    uint32_t       total = 0;
    uint32_t const rank  = block_rank(valid, &total);
    if (!valid) {
        return;
    }
    uint32_t const count = w.counters[C_records];
    uint32_t const salt  = *step * 2654435761u;  // a new permutation after every Adam step
    uint32_t const j     = w.rec_block[blockIdx.x] + rank;
    uint32_t const k     = uint32_t((uint64_t(j) * 2147483647ull + salt) % count);
    for (int c = 0; c < 9; ++c) {
        sx[9 * size_t(k) + c] = w.train_x[9 * size_t(i) + c];
    }
    for (int c = 0; c < 3; ++c) {
        sd[3 * size_t(k) + c] = w.train_dir[3 * size_t(i) + c];
    }
    sw[k] = w.train_w[i];
}

__global__ void reset_records(uint32_t* counters) {
    counters[C_records] = 0;
}

// Sums the wave's sample planes per pixel, in plane order, into the FP64
// accumulator (rgb per pixel).
__global__ void __launch_bounds__(256) resolve(Wave w, uint32_t planes, double* accum) {
    uint32_t const pixel = blockIdx.x * blockDim.x + threadIdx.x;
    if (pixel >= w.num_pixels) {
        return;
    }
    double s[3] = {0.0, 0.0, 0.0};
    for (uint32_t k = 0; k < planes; ++k) {
        Path const& p = w.paths[size_t(k) * w.num_pixels + pixel];
        for (int c = 0; c < 3; ++c) {
            s[c] += double(p.radiance[c]);
        }
    }
    for (int c = 0; c < 3; ++c) {
        accum[3 * size_t(pixel) + c] += s[c];
    }
}

// ---- filtered sensor ----

__device__ float filter_eval(Sensor_filter const& f, float d) {
    if (MI_FILTER_MITCHELL == f.kind) {
        // Mitchell::mitchell (sensor_mitchell.cpp), unclipped as there
        float const x  = fabsf(2.f * (d * f.radius_inv));
        float const xx = x * x;
        float const b = f.b, c = f.c;
        if (x > 1.f) {
            return ((((-b - 6.f * c) * xx) * x + (6.f * b + 30.f * c) * xx) + (-12.f * b - 48.f * c) * x +
                    (8.f * b + 24.f * c)) *
                   (1.f / 6.f);
        }
        return ((((12.f - 9.f * b) - 6.f * c) * xx) * x + ((-18.f + 12.f * b) + 6.f * c) * xx +
                (6.f - 2.f * b)) *
               (1.f / 6.f);
    }
    // Interpolated_function_1D_N<float, 30>::operator() (interpolated_function_1d.inl:168-178)
    float          x      = MI_FILTER_GAUSSIAN == f.kind ? d * d : fabsf(d);
    x                     = fminf(x, f.range_end);
    float const    o      = x * f.inverse_interval;
    uint32_t const offset = uint32_t(o);
    float const    t      = o - float(offset);
    uint32_t const next   = offset + 1 < 29u ? offset + 1 : 29u;
    return (1.f - t) * f.table[offset] + t * f.table[next];
}

// Gather form of Filtered_*::add_sample: image pixel (x, y) takes from every
// grid pixel within reach and every sample of the wave, in a fixed order,
// weight f(ox - k) f(oy - l) * clamped colour and the weight, where (k, l) is
// its offset from the sample's pixel and o = jitter - 0.5.
__global__ void __launch_bounds__(256) resolve_filtered(Wave w, uint32_t planes, double* accum) {
    uint32_t const pixel = blockIdx.x * blockDim.x + threadIdx.x;
    if (pixel >= uint32_t(w.width) * w.height) {
        return;
    }
    Sensor_filter const& f  = w.filter;
    int32_t const        bd = w.border;
    int32_t const        x  = int32_t(pixel % w.width);
    int32_t const        y  = int32_t(pixel / w.width);
    double               s[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t k = 0; k < planes; ++k) {
        for (int32_t l = -bd; l <= bd; ++l) {
            // the sample pixel lies l rows above: grid row y - l + bd
            uint32_t const gy = uint32_t(y - l + bd);
            for (int32_t kx = -bd; kx <= bd; ++kx) {
                uint32_t const gx = uint32_t(x - kx + bd);
                Path const&    p  = w.paths[size_t(k) * w.num_pixels + size_t(gy) * w.grid_w + gx];
                float const    wx = MI_FILTER_BOX == f.kind ? 1.f : filter_eval(f, (p.jitter[0] - 0.5f) - float(kx));
                float const    wy = MI_FILTER_BOX == f.kind ? 1.f : filter_eval(f, (p.jitter[1] - 0.5f) - float(l));
                float const    wt = wx * wy;
                float          c[3] = {p.radiance[0], p.radiance[1], p.radiance[2]};
                if (1 == f.clamp_mode) {
                    for (int i = 0; i < 3; ++i) {
                        c[i] = fminf(c[i], f.clamp_max[i]);
                    }
                } else if (2 == f.clamp_mode) {
                    float const mc = fmaxf(fmaxf(c[0], c[1]), c[2]);
                    if (mc > f.clamp_max[0]) {
                        float const r = f.clamp_max[0] / mc;
                        for (int i = 0; i < 3; ++i) {
                            c[i] = r * c[i];
                        }
                    }
                }
                for (int i = 0; i < 3; ++i) {
                    s[i] += double(wt * c[i]);
                }
                s[3] += double(wt);
            }
        }
    }
    for (int i = 0; i < 4; ++i) {
        accum[4 * size_t(pixel) + i] += s[i];
    }
}

// ---- energy table: sprout's procedure (ggx_integrate.cpp:17-45, 202-240) ----

__device__ float radical_inverse_vdc(uint32_t bits) {
    bits = (bits << 16) | (bits >> 16);
    bits = ((bits & 0x55555555u) << 1) | ((bits & 0xAAAAAAAAu) >> 1);
    bits = ((bits & 0x33333333u) << 2) | ((bits & 0xCCCCCCCCu) >> 2);
    bits = ((bits & 0x0F0F0F0Fu) << 4) | ((bits & 0xF0F0F0F0u) >> 4);
    bits = ((bits & 0x00FF00FFu) << 8) | ((bits & 0xFF00FF00u) >> 8);
    return float(bits) * 2.3283064365386963e-10f;
}

// One workgroup per table entry; the 1024 sample terms are summed in order
// by one lane, as the reference's loop does.
__global__ void __launch_bounds__(256) energy_table(float* table) {
    __shared__ float terms[1024];

    uint32_t const entry = blockIdx.x;
    uint32_t const ai    = entry / dev::E_size;
    uint32_t const ni    = entry % dev::E_size;

    // grid coordinates by float accumulation of the step, as generated
    float const step  = 1.f / float(dev::E_size - 1);
    float       alpha = 0.f;
    for (uint32_t k = 0; k < ai; ++k) {
        alpha += step;
    }
    float n_dot_wo = 0.f;
    for (uint32_t k = 0; k < ni; ++k) {
        n_dot_wo += step;
    }

    if (alpha < dev::Min_rough * dev::Min_rough) {
        if (0 == threadIdx.x) {
            table[entry] = 1.f;
        }
        return;
    }

    n_dot_wo    = dev::clampf(n_dot_wo, 0.f, 1.f);
    V3 const wo = v3(sqrtf(1.f - n_dot_wo * n_dot_wo), 0.f, n_dot_wo);

    dev::Material_closure m;
    m.layer.t  = v3(1.f, 0.f, 0.f);
    m.layer.b  = v3(0.f, 1.f, 0.f);
    m.layer.n  = v3(0.f, 0.f, 1.f);
    m.geo_n    = m.layer.n;
    m.wo       = wo;
    m.alpha    = alpha;
    m.metallic = 1.f;
    m.e_table  = nullptr;
    for (int c = 0; c < 3; ++c) {
        m.f0[c]     = 1.f;
        m.albedo[c] = 0.f;
    }

    for (uint32_t i = threadIdx.x; i < 1024; i += blockDim.x) {
        dev::Bxdf_sample s;
        float const n_dot_wi = m.ggx_reflect(n_dot_wo, float(i) / 1024.f, radical_inverse_vdc(i), s);
        terms[i]             = (n_dot_wi * s.refl[0]) / s.pdf;
    }
    __syncthreads();
    if (0 == threadIdx.x) {
        float accum = 0.f;
        for (uint32_t i = 0; i < 1024; ++i) {
            accum += terms[i];
        }
        table[entry] = accum / 1024.f;
    }
}

__device__ dev::Material_closure batch_closure(uint32_t i, float const* params, float const* frames,
                                               float const* wo, float const* table) {
    float const* f = frames + 12 * size_t(i);
    float const* q = params + 6 * size_t(i);
    dev::Layer   l;
    l.t = v3(f[0], f[1], f[2]);
    l.b = v3(f[3], f[4], f[5]);
    l.n = v3(f[6], f[7], f[8]);
    dev::Material_closure m;
    m.set(l, v3(f[9], f[10], f[11]), v3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), q, q[3], q[4],
          q[5], 1.f, table);
    return m;
}

__global__ void bsdf_evaluate_batch(uint32_t n, float const* params, float const* frames,
                                    float const* wo, float const* wi, float const* table,
                                    float* out) {
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    dev::Material_closure const m = batch_closure(i, params, frames, wo, table);
    m.evaluate(v3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]), out + 4 * size_t(i));
}

__global__ void bsdf_sample_batch(uint32_t n, float const* params, float const* frames,
                                  float const* wo, float const* xi, float const* table,
                                  float* out) {
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    dev::Material_closure const m = batch_closure(i, params, frames, wo, table);
    dev::Bxdf_sample            s;
    m.sample(xi[3 * i], xi[3 * i + 1], xi[3 * i + 2], s);
    float* o = out + 8 * size_t(i);
    if (!(s.pdf > 0.f)) {
        for (int k = 0; k < 8; ++k) {
            o[k] = 0.f;
        }
        return;
    }
    o[0] = s.wi.x;
    o[1] = s.wi.y;
    o[2] = s.wi.z;
    o[3] = s.refl[0];
    o[4] = s.refl[1];
    o[5] = s.refl[2];
    o[6] = s.pdf;
    o[7] = __uint_as_float(s.type);
}

// Batch tracking along rays through a scene's medium (each ray clipped to
// the box as the renderer clips it), one generator per ray started at
// (seed, index): track_segment -> {event, t, w rgb}, or with
// MI_MEDIUM_TRANSMITTANCE transmittance -> {transmitted, 0, tr rgb}.
__global__ void medium_batch(dev::Medium_view m, mi_ray const* rays, uint32_t depth, uint64_t seed,
                             uint32_t flags, uint32_t n, float* out) {
    bool const transmittance = 0 != (flags & MI_MEDIUM_TRANSMITTANCE);
    bool const nearest       = 0 != (flags & MI_MEDIUM_NEAREST);
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    mi_ray const r = rays[i];
    V3 const     o = v3(r.origin[0], r.origin[1], r.origin[2]);
    V3 const     d = v3(r.direction[0], r.direction[1], r.direction[2]);
    Rng          rng;
    rng.start(seed, i);
    float res[5] = {0.f, 0.f, 1.f, 1.f, 1.f};
    float t0, t1;
    if (dev::box_overlap(m, o, d, r.min_t, r.max_t, t0, t1)) {
        if (transmittance) {
            res[0] = dev::transmittance(m, o, d, t0, t1, depth, nearest, rng, res + 2) ? 1.f : 0.f;
        } else {
            float t_out = 0.f;
            res[0]      = float(dev::track_segment(m, o, d, t0, t1, depth, nearest, rng, t_out, res + 2));
            res[1]      = t_out;
        }
    } else if (transmittance) {
        res[0] = 1.f;
    }
    for (int k = 0; k < 5; ++k) {
        out[5 * size_t(i) + k] = res[k];
    }
}

__global__ void light_tree_sample_batch(dev::Light_tree_view t, float const* points,
                                        float const* normals, float const* randoms, uint32_t split,
                                        uint32_t n, uint32_t* picks, uint32_t* counts) {
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    V3 const       p  = v3(points[3 * i], points[3 * i + 1], points[3 * i + 2]);
    V3 const       nn = v3(normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]);
    uint32_t*      o  = picks + 2 * size_t(i) * dev::Max_lights;
    uint32_t const c  = dev::for_each_light(t, p, nn, 0 != (split & MI_LIGHT_TOTAL_SPHERE), randoms[i],
                                            0 != (split & MI_LIGHT_SPLIT),
                                            [o](uint32_t k, dev::Light_pick const pick) {
                                                o[2 * k]     = pick.light;
                                                o[2 * k + 1] = __float_as_uint(pick.pdf);
                                            });
    // the count the shading kernels reserve shadow rays by
    counts[i] = c == dev::count_lights(t, p, 0 != (split & MI_LIGHT_SPLIT)) ? c : dev::Null;
}

__global__ void light_tree_pdf_batch(dev::Light_tree_view t, float const* points,
                                     float const* normals, uint32_t const* lights, uint32_t split,
                                     uint32_t n, float* pdfs) {
    uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    V3 const p  = v3(points[3 * i], points[3 * i + 1], points[3 * i + 2]);
    V3 const nn = v3(normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]);
    pdfs[i]     = dev::light_pdf(t, p, nn, 0 != (split & MI_LIGHT_TOTAL_SPHERE),
                                 0 != (split & MI_LIGHT_SPLIT), lights[i]);
}

uint32_t blocks(uint64_t n, uint32_t b) {
    uint64_t const g = (n + b - 1) / b;
    return uint32_t(g > 0 ? g : 1);
}

bool is_light(mi_material const& m) {
    return MI_MATERIAL_LIGHT == m.kind;
}

}  // namespace

int32_t make_energy_table(float** table) {
    float* t = nullptr;
    if (hipError_t const e = hipMalloc(&t, dev::E_size * dev::E_size * sizeof(float)); e) {
        return fail(MI_ERR_HIP, "energy table: %s", hipGetErrorString(e));
    }
    hipLaunchKernelGGL(energy_table, dim3(dev::E_size * dev::E_size), dim3(256), 0, nullptr, t);
    if (hipError_t const e = hipDeviceSynchronize(); e) {
        (void)hipFree(t);
        return fail(MI_ERR_HIP, "energy table: %s", hipGetErrorString(e));
    }
    *table = t;
    return MI_OK;
}

int32_t bsdf_batch(bool sample, uint32_t n, float const* params, float const* frames,
                   float const* wo, float const* wi_or_xi, float const* table, float* out,
                   hipStream_t stream) {
    if (0 == n) {
        return MI_OK;
    }
    if (sample) {
        hipLaunchKernelGGL(bsdf_sample_batch, dim3(blocks(n, 256)), dim3(256), 0, stream, n,
                           params, frames, wo, wi_or_xi, table, out);
    } else {
        hipLaunchKernelGGL(bsdf_evaluate_batch, dim3(blocks(n, 256)), dim3(256), 0, stream, n,
                           params, frames, wo, wi_or_xi, table, out);
    }
    if (hipError_t const e = hipGetLastError(); e) {
        return fail(MI_ERR_HIP, "bsdf batch: %s", hipGetErrorString(e));
    }
    return MI_OK;
}

namespace {

dev::Light_tree_view tree_view(RenderScene const& scene) {
    dev::Light_tree_view t;
    t.nodes     = scene.light_nodes;
    t.middles   = scene.light_middles;
    t.mapping   = scene.light_mapping;
    t.orders    = scene.light_orders;
    t.lights    = scene.light_infos;
    t.num_nodes = scene.num_lights ? scene.num_light_nodes : 0;
    return t;
}

}  // namespace

int32_t light_tree_queries(RenderScene const& scene, bool sample, float const* points,
                           float const* normals, void const* arg, uint32_t split, uint32_t n,
                           void* out0, uint32_t* out1, hipStream_t stream) {
    if (0 == n) {
        return MI_OK;
    }
    dev::Light_tree_view const t = tree_view(scene);
    if (sample) {
        hipLaunchKernelGGL(light_tree_sample_batch, dim3(blocks(n, 128)), dim3(128), 0, stream, t,
                           points, normals, static_cast<float const*>(arg), split, n,
                           static_cast<uint32_t*>(out0), out1);
    } else {
        hipLaunchKernelGGL(light_tree_pdf_batch, dim3(blocks(n, 128)), dim3(128), 0, stream, t,
                           points, normals, static_cast<uint32_t const*>(arg), split, n,
                           static_cast<float*>(out0));
    }
    if (hipError_t const e = hipGetLastError(); e) {
        return fail(MI_ERR_HIP, "light tree query: %s", hipGetErrorString(e));
    }
    return MI_OK;
}

int32_t medium_queries(RenderScene const& scene, mi_ray const* rays, uint32_t depth, uint64_t seed,
                       uint32_t flags, uint32_t n, float* out, hipStream_t stream) {
    if (!scene.medium.present) {
        return fail(MI_ERR_ARGS, "mi_medium_track: the scene has no medium");
    }
    if (0 == n) {
        return MI_OK;
    }
    hipLaunchKernelGGL(medium_batch, dim3(blocks(n, 128)), dim3(128), 0, stream, scene.medium, rays,
                       depth, seed, flags, n, out);
    if (hipError_t const e = hipGetLastError(); e) {
        return fail(MI_ERR_HIP, "mi_medium_track: %s", hipGetErrorString(e));
    }
    return MI_OK;
}

int32_t download_light_tree(RenderScene const& scene, mi_light_node* nodes, uint32_t* middles,
                            uint32_t* mapping, uint32_t* orders, void* infos) {
    size_t const nn = scene.num_lights ? scene.num_light_nodes : 0;
    size_t const nl = scene.num_lights;
    hipError_t   e  = hipSuccess;
    if (nodes && nn) e = hipMemcpy(nodes, scene.light_nodes, nn * sizeof(Light_node), hipMemcpyDeviceToHost);
    if (!e && middles && nn) e = hipMemcpy(middles, scene.light_middles, nn * sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (!e && mapping && nl) e = hipMemcpy(mapping, scene.light_mapping, nl * sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (!e && orders && nl) e = hipMemcpy(orders, scene.light_orders, nl * sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (!e && infos && nl) e = hipMemcpy(infos, scene.light_infos, nl * sizeof(Light_info), hipMemcpyDeviceToHost);
    if (e) {
        return fail(MI_ERR_HIP, "light tree download: %s", hipGetErrorString(e));
    }
    return MI_OK;
}

void RenderScene::release() {
    bvh.release();
    (void)hipFree(positions);
    (void)hipFree(normals);
    (void)hipFree(indices);
    (void)hipFree(tri_material);
    (void)hipFree(materials);
    (void)hipFree(light_nodes);
    (void)hipFree(light_middles);
    (void)hipFree(light_mapping);
    (void)hipFree(light_orders);
    (void)hipFree(light_infos);
    (void)hipFree(tri_light);
    (void)hipFree(tri_local);
    (void)hipFree(part_infos);
    (void)hipFree(part_nodes);
    (void)hipFree(part_middles);
    (void)hipFree(part_cdf_offset);
    (void)hipFree(part_cdfs);
    (void)hipFree(part_mapping);
    (void)hipFree(part_orders);
    (void)hipFree(part_lights);
    (void)hipFree(emit_tris);
    (void)hipFree(part_radiance);
    emitter_work.release();
    release_medium();
    tlas.release();
    for (DeviceBvh& t : mesh_trees) {
        t.release();
    }
    for (float* p : mesh_normals) {
        (void)hipFree(p);
    }
    mesh_normals.clear();
    for (void* p : mesh_buffers) {
        (void)hipFree(p);
    }
    (void)hipFree(mesh_geo);
    (void)hipFree(inst_o2w);
    (void)hipFree(inst_material);
    *this = RenderScene();
}

void RenderScene::release_medium() {
    (void)hipFree(medium_density);
    (void)hipFree(medium_nodes);
    (void)hipFree(medium_data);
    medium_density = nullptr;
    medium_nodes   = nullptr;
    medium_data    = nullptr;
    medium         = dev::Medium_view{};
}

constexpr double Max_optical_depth = 1.0e4;

namespace {

// Unit copies of per-vertex normals on the device (reusing *d when set).
int32_t upload_unit_normals(float const* normals, uint32_t num_vertices, float** d,
                            char const* who) {
    std::vector<float> unit(size_t(num_vertices) * 3);
    for (uint32_t v = 0; v < num_vertices; ++v) {
        float const* n  = normals + 3 * size_t(v);
        float const  l2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
        if (!(l2 > 0.f) || !std::isfinite(l2)) {
            return fail(MI_ERR_ARGS, "%s: vertex %u has no direction", who, v);
        }
        float const il = 1.f / std::sqrt(l2);
        for (int k = 0; k < 3; ++k) {
            unit[3 * size_t(v) + k] = n[k] * il;
        }
    }
    float*     p = *d;
    hipError_t e = hipSuccess;
    if (!p) {
        e = hipMalloc(&p, unit.size() * sizeof(float));
    }
    if (!e) e = hipMemcpy(p, unit.data(), unit.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e) {
        if (p != *d) {
            (void)hipFree(p);
        }
        return fail(MI_ERR_HIP, "%s: %s", who, hipGetErrorString(e));
    }
    *d = p;
    return MI_OK;
}

}  // namespace

int32_t set_mesh_normals(RenderScene& s, uint32_t mesh, float const* normals,
                         uint32_t num_vertices) {
    if (!s.instanced || mesh >= s.mesh_geo_host.size()) {
        return fail(MI_ERR_ARGS, "mi_scene_set_mesh_normals: mesh %u of %zu", mesh,
                    s.mesh_geo_host.size());
    }
    if (normals && num_vertices != s.mesh_vertices[mesh]) {
        return fail(MI_ERR_ARGS, "mi_scene_set_mesh_normals: %u normals for %u vertices",
                    num_vertices, s.mesh_vertices[mesh]);
    }
    float* d = s.mesh_normals[mesh];
    if (normals) {
        if (int32_t const st = upload_unit_normals(normals, num_vertices, &d,
                                                   "mi_scene_set_mesh_normals");
            st) {
            return st;
        }
    } else {
        (void)hipFree(d);
        d = nullptr;
    }
    s.mesh_normals[mesh]          = d;
    s.mesh_geo_host[mesh].normals = d;
    hipError_t const e = hipMemcpy(s.mesh_geo + mesh, &s.mesh_geo_host[mesh], sizeof(Mesh_geometry),
                                   hipMemcpyHostToDevice);
    return e ? fail(MI_ERR_HIP, "mi_scene_set_mesh_normals: %s", hipGetErrorString(e)) : MI_OK;
}

int32_t set_normals(RenderScene& s, float const* normals, uint32_t num_vertices) {
    if (!normals) {
        (void)hipFree(s.normals);
        s.normals = nullptr;
        return MI_OK;
    }
    if (num_vertices != s.num_vertices) {
        return fail(MI_ERR_ARGS, "mi_scene_set_normals: %u normals for %u vertices", num_vertices,
                    s.num_vertices);
    }
    // sprout keeps a vertex's tangent frame as a unit quaternion
    // (triangle_bvh_indexed_data.inl:136-152), so its normals are unit length
    return upload_unit_normals(normals, num_vertices, &s.normals, "mi_scene_set_normals");
}

int32_t set_medium(RenderScene& scene, float const* density, uint32_t const dims[3],
                   float const box_lo[3], float const box_hi[3], float const cc_a[3],
                   float const cc_s[3], float g, uint32_t majorants) {
    scene.release_medium();
    if (majorants > MI_MAJORANTS_FLAT) {
        return fail(MI_ERR_ARGS, "medium majorants %u unknown", majorants);
    }
    if (!density) {
        return MI_OK;
    }
    for (int k = 0; k < 3; ++k) {
        if (0 == dims[k] || dims[k] > 4096) {
            return fail(MI_ERR_ARGS, "medium grid dimension %u out of [1, 4096]", dims[k]);
        }
        if (!(box_hi[k] > box_lo[k]) || !std::isfinite(box_lo[k]) || !std::isfinite(box_hi[k])) {
            return fail(MI_ERR_ARGS, "medium box is empty or not finite");
        }
        if (!(cc_a[k] >= 0.f) || !(cc_s[k] >= 0.f) || !std::isfinite(cc_a[k]) ||
            !std::isfinite(cc_s[k])) {
            return fail(MI_ERR_ARGS, "medium coefficients must be finite and non-negative");
        }
    }
    if (!(g > -1.f && g < 1.f)) {
        return fail(MI_ERR_ARGS, "phase anisotropy g = %g outside (-1, 1)", double(g));
    }
    size_t const n     = size_t(dims[0]) * dims[1] * dims[2];
    float        max_d = 0.f;
    for (size_t i = 0; i < n; ++i) {
        if (!(density[i] >= 0.f) || !std::isfinite(density[i])) {
            return fail(MI_ERR_ARGS, "medium density %zu is negative or not finite", i);
        }
        max_d = std::max(max_d, density[i]);
    }
    // delta tracking takes about (majorant x length) steps per ray: bound the
    // optical depth of the box diagonal so that every tracking loop ends soon
    double diag2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        double const e = double(box_hi[k]) - double(box_lo[k]);
        diag2 += e * e;
    }
    double const mt_max = double(max_d) * (double(std::max({cc_a[0], cc_a[1], cc_a[2]})) +
                                           double(std::max({cc_s[0], cc_s[1], cc_s[2]})));
    if (mt_max * std::sqrt(diag2) > Max_optical_depth) {
        return fail(MI_ERR_ARGS, "medium optical depth %g across the box exceeds %g",
                    mt_max * std::sqrt(diag2), Max_optical_depth);
    }
    int32_t const idims[3] = {int32_t(dims[0]), int32_t(dims[1]), int32_t(dims[2])};
    HostGridtree  tree;
    build_gridtree(density, idims, cc_a, cc_s, tree, MI_MAJORANTS_FLAT == majorants);

    hipError_t e = hipMalloc(&scene.medium_density, n * sizeof(float));
    if (!e) e = hipMalloc(&scene.medium_nodes, tree.nodes.size() * sizeof(uint32_t));
    if (!e) e = hipMalloc(&scene.medium_data, std::max<size_t>(tree.data.size(), 1) * sizeof(Medium_cm));
    if (!e) e = hipMemcpy(scene.medium_density, density, n * sizeof(float), hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(scene.medium_nodes, tree.nodes.data(), tree.nodes.size() * sizeof(uint32_t),
                          hipMemcpyHostToDevice);
    if (!e && !tree.data.empty()) {
        e = hipMemcpy(scene.medium_data, tree.data.data(), tree.data.size() * sizeof(Medium_cm),
                      hipMemcpyHostToDevice);
    }
    if (e) {
        scene.release_medium();
        return fail(MI_ERR_HIP, "medium upload: %s", hipGetErrorString(e));
    }
    dev::Medium_view& m = scene.medium;
    m.density           = scene.medium_density;
    for (int k = 0; k < 3; ++k) {
        m.dims[k]      = idims[k];
        m.inv_dims[k]  = 1.f / float(idims[k]);
        m.num_cells[k] = uint32_t(tree.num_cells[k]);
        m.box_lo[k]    = box_lo[k];
        m.iext[k]      = 1.f / (box_hi[k] - box_lo[k]);
        m.cc_a[k]      = cc_a[k];
        m.cc_s[k]      = cc_s[k];
    }
    m.nodes   = scene.medium_nodes;
    m.data    = scene.medium_data;
    m.g       = g;
    m.present = 1;
    return MI_OK;
}

int32_t create_scene(float const* positions, uint32_t num_vertices, uint32_t const* indices,
                     uint32_t num_triangles, uint32_t const* tri_material,
                     mi_material const* materials, uint32_t num_materials,
                     mi_bvh_settings const& settings, RenderScene& out) {
    for (uint64_t i = 0; i < 3ull * num_triangles; ++i) {
        if (indices[i] >= num_vertices) {
            return fail(MI_ERR_ARGS, "mi_scene_create: index %llu is %u, of %u vertices",
                        (unsigned long long)i, indices[i], num_vertices);
        }
    }
    for (uint32_t t = 0; t < num_triangles; ++t) {
        if (tri_material[t] >= num_materials) {
            return fail(MI_ERR_ARGS, "mi_scene_create: triangle %u has material %u of %u", t,
                        tri_material[t], num_materials);
        }
    }
    std::vector<Render_material> mats(num_materials);
    for (uint32_t m = 0; m < num_materials; ++m) {
        mi_material const& in = materials[m];
        if (in.kind > MI_MATERIAL_LIGHT) {
            return fail(MI_ERR_ARGS, "mi_scene_create: material %u has kind %u", m, in.kind);
        }
        Render_material& r = mats[m];
        std::memset(&r, 0, sizeof(r));
        for (int c = 0; c < 3; ++c) {
            r.color[c]    = in.color[c];
            r.emission[c] = is_light(in) ? in.emission[c] : 0.f;
        }
        r.roughness = in.roughness;
        r.metallic  = in.metallic;
        r.ior       = in.ior;
        r.kind      = in.kind;
    }

    // Emissive parts: the triangles of one emissive material, in the order
    // the material first appears (a mesh part, triangle_mesh.cpp:67-247).
    // Each part is one light of the scene tree, with the box, cone and
    // power Scene::prop_prepare_sampling gives it (scene.cpp:440-506,
    // light.cpp:38-48), and has a Primitive_tree over its triangles
    // (power = area, one-sided).
    std::vector<int32_t>               part_of(num_materials, -1);
    std::vector<std::vector<uint32_t>> part_tris;
    std::vector<uint32_t>              tri_light(num_triangles, dev::Null);
    std::vector<uint32_t>              tri_local(num_triangles, dev::Null);
    std::vector<Light_info>            tri_info(num_triangles);
    for (uint32_t t = 0; t < num_triangles; ++t) {
        mi_material const& mat = materials[tri_material[t]];
        if (!is_light(mat)) {
            continue;
        }
        float const* a = positions + 3 * size_t(indices[3 * size_t(t) + 0]);
        float const* b = positions + 3 * size_t(indices[3 * size_t(t) + 1]);
        float const* c = positions + 3 * size_t(indices[3 * size_t(t) + 2]);
        float const  radiance = std::max(std::max(mat.emission[0], mat.emission[1]), mat.emission[2]);
        lb::Tri_light tl;
        if (!lb::triangle_light(a, b, c, tl) || !(radiance > 0.f)) {
            continue;
        }
        Light_info& li = tri_info[t];
        for (int k = 0; k < 3; ++k) {
            li.center[k] = tl.center[k];
            li.cone[k]   = tl.n[k];
        }
        li.radius    = tl.radius;
        li.cone[3]   = 1.f;
        li.power     = tl.area;  // the part tree's power (no emission map)
        li.two_sided = 0;
        li.triangle  = t;
        li.area      = tl.area;
        uint32_t const m = tri_material[t];
        if (part_of[m] < 0) {
            part_of[m] = int32_t(part_tris.size());
            part_tris.emplace_back();
        }
        tri_light[t] = uint32_t(part_of[m]);
        tri_local[t] = uint32_t(part_tris[part_of[m]].size());
        part_tris[part_of[m]].push_back(t);
    }

    uint32_t const          num_parts = uint32_t(part_tris.size());
    std::vector<Light_info> infos(num_parts);
    std::vector<float>      bmin, bmax, cones, powers;
    std::vector<uint8_t>    two_sided(num_parts, 0);
    // the part pools, laid out as the device builder writes them: part p
    // with its lights from first_p keeps its nodes from 2 first_p - p (room
    // for 2 n_p - 1) and its leaf distributions from 2 first_p (room for 2 n_p)
    uint32_t num_emitters = 0;
    for (auto const& tris : part_tris) {
        num_emitters += uint32_t(tris.size());
    }
    std::vector<dev::Part_info> pinfo(num_parts);
    std::vector<Light_node>     pnodes(std::max<size_t>(1, 2 * size_t(num_emitters)));
    std::vector<uint32_t>       pmiddles(pnodes.size(), 0), pcdf_off(pnodes.size(), dev::Null);
    std::vector<uint32_t>       pmapping, porders, emit_tris, part_first(1, 0);
    std::vector<float>          pcdfs(std::max<size_t>(1, 2 * size_t(num_emitters)), 0.f);
    std::vector<float>          part_radiance(num_parts);
    std::vector<Light_info>     plights;
    for (uint32_t p = 0; p < num_parts; ++p) {
        std::vector<uint32_t> const& tris = part_tris[p];
        uint32_t const               n    = uint32_t(tris.size());
        std::vector<float>           tmin(3 * size_t(n)), tmax(3 * size_t(n)), tcone(4 * size_t(n)), tpow(n);
        auto const tri = [&](uint32_t i) {
            float const*  a = positions + 3 * size_t(indices[3 * size_t(tris[i]) + 0]);
            float const*  b = positions + 3 * size_t(indices[3 * size_t(tris[i]) + 1]);
            float const*  c = positions + 3 * size_t(indices[3 * size_t(tris[i]) + 2]);
            lb::Tri_light t;
            lb::triangle_light(a, b, c, t);
            return t;
        };
        for (uint32_t i = 0; i < n; ++i) {
            lb::Tri_light const t = tri(i);
            for (int k = 0; k < 3; ++k) {
                tmin[3 * i + k] = t.lo[k];
                tmax[3 * i + k] = t.hi[k];
            }
            for (int k = 0; k < 4; ++k) {
                tcone[4 * i + k] = tri_info[tris[i]].cone[k];
            }
            tpow[i] = tri_info[tris[i]].power;
        }
        // Part::init: the dominant axis and the widest angle from it
        lb::Part_light pl;
        lb::part_light(n, tri, pl);
        float const* lo        = pl.lo;
        float const* hi        = pl.hi;
        float const  total     = pl.total;
        float const* part_cone = pl.cone;

        HostLightTree tree;
        build_part_tree(tmin.data(), tmax.data(), tcone.data(), tpow.data(), n, part_cone, tree);
        dev::Part_info& pi      = pinfo[p];
        uint32_t const  first_p = uint32_t(plights.size());
        pi.node_offset          = 2 * first_p - p;
        pi.light_offset         = first_p;
        pi.num_nodes            = uint32_t(tree.nodes.size());
        pi.num_lights           = n;
        uint32_t const cdf_base = 2 * first_p;
        for (uint32_t i = 0; i < tree.nodes.size(); ++i) {
            pnodes[pi.node_offset + i]   = tree.nodes[i];
            pmiddles[pi.node_offset + i] = tree.middles[i];
            pcdf_off[pi.node_offset + i] =
                tree.cdf_offset[i] == 0xFFFFFFFFu ? dev::Null : cdf_base + tree.cdf_offset[i];
        }
        std::copy(tree.cdfs.begin(), tree.cdfs.end(), pcdfs.begin() + cdf_base);
        for (uint32_t i = 0; i < n; ++i) {
            pmapping.push_back(tree.light_mapping[i]);
            porders.push_back(tree.light_orders[i]);
            plights.push_back(tri_info[tris[i]]);
            emit_tris.push_back(tris[i]);
        }
        part_first.push_back(uint32_t(plights.size()));

        // the scene light of the part
        float const radiance = std::max(std::max(materials[tri_material[tris[0]]].emission[0],
                                                 materials[tri_material[tris[0]]].emission[1]),
                                        materials[tri_material[tris[0]]].emission[2]);
        Light_info& L = infos[p];
        for (int k = 0; k < 3; ++k) {
            bmin.push_back(lo[k]);
            bmax.push_back(hi[k]);
            L.center[k] = 0.5f * (lo[k] + hi[k]);
        }
        float const ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
        L.radius    = 0.5f * std::sqrt((ext[0] * ext[0] + ext[1] * ext[1]) + ext[2] * ext[2]);
        for (int k = 0; k < 4; ++k) {
            L.cone[k] = part_cone[k];
            cones.push_back(part_cone[k]);
        }
        L.power     = total * radiance;
        L.two_sided = 0;
        L.triangle  = p;
        L.area      = total;
        powers.push_back(L.power);
        part_radiance[p] = radiance;
    }
    HostLightTree tree;
    build_light_tree(bmin.data(), bmax.data(), cones.data(), powers.data(), two_sided.data(),
                     num_parts, tree);

    RenderScene s;
    s.num_vertices  = num_vertices;
    s.num_triangles = num_triangles;
    s.num_materials = num_materials;
    s.num_lights    = uint32_t(infos.size());
    s.num_light_nodes = uint32_t(tree.nodes.size());
    for (int k = 0; k < 3; ++k) {
        s.box_lo[k] = positions[k];
        s.box_hi[k] = positions[k];
    }
    for (uint32_t v = 0; v < num_vertices; ++v) {
        for (int k = 0; k < 3; ++k) {
            s.box_lo[k] = std::min(s.box_lo[k], positions[3 * size_t(v) + k]);
            s.box_hi[k] = std::max(s.box_hi[k], positions[3 * size_t(v) + k]);
        }
    }

    size_t const nl = std::max<size_t>(1, infos.size());
    size_t const nn = std::max<size_t>(1, 2 * infos.size());  // room for any tree over the parts
    hipError_t   e  = hipMalloc(&s.positions, size_t(num_vertices) * 3 * sizeof(float));
    if (!e) e = hipMalloc(&s.indices, size_t(num_triangles) * 3 * sizeof(uint32_t));
    if (!e) e = hipMalloc(&s.tri_material, size_t(num_triangles) * sizeof(uint32_t));
    if (!e) e = hipMalloc(&s.materials, std::max<size_t>(1, num_materials) * sizeof(Render_material));
    if (!e) e = hipMalloc(&s.light_nodes, nn * sizeof(Light_node));
    if (!e) e = hipMalloc(&s.light_middles, nn * sizeof(uint32_t));
    if (!e) e = hipMalloc(&s.light_mapping, nl * sizeof(uint32_t));
    if (!e) e = hipMalloc(&s.light_orders, nl * sizeof(uint32_t));
    if (!e) e = hipMalloc(&s.light_infos, nl * sizeof(Light_info));
    if (!e) e = hipMalloc(&s.tri_light, size_t(num_triangles) * sizeof(uint32_t));
    if (!e) e = hipMemcpy(s.positions, positions, size_t(num_vertices) * 3 * sizeof(float), hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(s.indices, indices, size_t(num_triangles) * 3 * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(s.tri_material, tri_material, size_t(num_triangles) * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(s.materials, mats.data(), num_materials * sizeof(Render_material), hipMemcpyHostToDevice);
    if (!e && !infos.empty()) {
        e = hipMemcpy(s.light_nodes, tree.nodes.data(), tree.nodes.size() * sizeof(Light_node), hipMemcpyHostToDevice);
        if (!e) e = hipMemcpy(s.light_middles, tree.middles.data(), tree.middles.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
        if (!e) e = hipMemcpy(s.light_mapping, tree.light_mapping.data(), infos.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
        if (!e) e = hipMemcpy(s.light_orders, tree.light_orders.data(), infos.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
        if (!e) e = hipMemcpy(s.light_infos, infos.data(), infos.size() * sizeof(Light_info), hipMemcpyHostToDevice);
    }
    if (!e) e = hipMemcpy(s.tri_light, tri_light.data(), size_t(num_triangles) * sizeof(uint32_t), hipMemcpyHostToDevice);
    // the part pools
    auto upload = [&e](auto*& dst, auto const& v) {
        using T = std::remove_reference_t<decltype(v[0])>;
        size_t const bytes = std::max<size_t>(1, v.size()) * sizeof(T);
        if (!e) e = hipMalloc(reinterpret_cast<void**>(&dst), bytes);
        if (!e && !v.empty()) e = hipMemcpy(dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    };
    upload(s.part_infos, pinfo);
    upload(s.part_nodes, pnodes);
    upload(s.part_middles, pmiddles);
    upload(s.part_cdf_offset, pcdf_off);
    upload(s.part_cdfs, pcdfs);
    upload(s.part_mapping, pmapping);
    upload(s.part_orders, porders);
    upload(s.part_lights, plights);
    upload(s.tri_local, tri_local);
    upload(s.emit_tris, emit_tris);
    upload(s.part_radiance, part_radiance);
    s.part_first     = part_first;
    s.part_info_host = pinfo;
    s.num_emitters   = num_emitters;
    s.bvh_settings   = settings;
    if (e) {
        s.release();
        return fail(MI_ERR_HIP, "mi_scene_create: %s", hipGetErrorString(e));
    }
    if (int32_t const st = build_bvh(s.positions, num_vertices, 3, s.indices, num_triangles,
                                     settings, nullptr, s.bvh);
        st) {
        s.release();
        return st;
    }
    out = s;
    return MI_OK;
}

namespace {

// Interpolated_function_1D_N<float, 30>::operator() on the host
float table_eval(Sensor_filter const& f, float x) {
    x                     = std::min(x, f.range_end);
    float const    o      = x * f.inverse_interval;
    uint32_t const offset = uint32_t(o);
    float const    t      = o - float(offset);
    uint32_t const next   = std::min(offset + 1, 29u);
    return (1.f - t) * f.table[offset] + t * f.table[next];
}

}  // namespace

// The tables as sprout's constructors fill and normalize them:
// Interpolated_function_1D_N (interpolated_function_1d.inl:124-137) over
// [0, r^2] (Gaussian, sensor_gaussian.cpp) or [0, r] (Blackman,
// sensor_blackman.cpp), scaled by 1 / the 64-sample midpoint integral of the
// unscaled filter over [-r, r].
int32_t make_sensor_filter(mi_sensor const& in, Sensor_filter& f) {
    std::memset(&f, 0, sizeof(f));
    if (in.filter > MI_FILTER_MITCHELL) {
        return fail(MI_ERR_ARGS, "sensor: filter %u", in.filter);
    }
    if (MI_FILTER_BOX != in.filter && !(in.radius > 0.f && in.radius <= 2.f)) {
        return fail(MI_ERR_ARGS, "sensor: radius %g outside (0, 2]", double(in.radius));
    }
    if (MI_FILTER_GAUSSIAN == in.filter && !(in.alpha > 0.f && std::isfinite(in.alpha))) {
        return fail(MI_ERR_ARGS, "sensor: Gaussian alpha %g", double(in.alpha));
    }
    if (MI_FILTER_MITCHELL == in.filter && !(std::isfinite(in.b) && std::isfinite(in.c))) {
        return fail(MI_ERR_ARGS, "sensor: Mitchell b %g, c %g", double(in.b), double(in.c));
    }
    f.kind       = in.filter;
    f.border     = MI_FILTER_BOX == in.filter ? 0 : (in.radius <= 1.f ? 1 : 2);
    f.radius_inv = MI_FILTER_BOX == in.filter ? 0.f : 1.f / in.radius;
    f.b          = in.b;
    f.c          = in.c;
    if (in.clamp_max[0] >= 0.f) {
        f.clamp_mode = in.clamp_max[1] < 0.f ? 2u : 1u;
        for (int i = 0; i < 3; ++i) {
            f.clamp_max[i] = in.clamp_max[i];
        }
    }
    if (MI_FILTER_GAUSSIAN == in.filter || MI_FILTER_BLACKMAN == in.filter) {
        bool const  gauss    = MI_FILTER_GAUSSIAN == in.filter;
        float const r        = in.radius;
        float const end      = gauss ? r * r : r;
        float const interval = end / 29.f;
        f.range_end          = end;
        f.inverse_interval   = 1.f / interval;
        float s              = 0.f;
        for (int i = 0; i < 30; ++i, s += interval) {
            if (gauss) {
                float const e = std::exp(-in.alpha * (r * r));
                f.table[i]    = std::max(0.f, std::exp(-in.alpha * s) - e);
            } else {
                float const a0 = 0.35875f, a1 = 0.48829f, a2 = 0.14128f, a3 = 0.01168f;
                float const x  = (3.14159265358979323846f * (s + r)) / r;
                f.table[i]     = ((a0 - a1 * std::cos(x)) + a2 * std::cos(2.f * x)) - a3 * std::cos(3.f * x);
            }
        }
        float const step = r / 64.f;
        float       m    = 0.5f * step;
        float       sum  = 0.f;
        for (int i = 0; i < 64; ++i, m += step) {
            sum += table_eval(f, gauss ? m * m : m) * step;
        }
        float const scale = 1.f / (sum + sum);
        for (float& v : f.table) {
            v *= scale;
        }
    }
    return MI_OK;
}

namespace {

// Every emissive triangle (part order) as a light: the part trees' inputs
// and the part's light records, as create_scene forms them.
__global__ void emitter_lights(float const* pos, uint32_t const* idx, uint32_t const* emit_tris,
                               uint32_t n, Emitter_work w, Light_info* plights) {
    uint32_t const j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) {
        return;
    }
    uint32_t const t = emit_tris[j];
    lb::Tri_light  tl;
    if (!lb::triangle_light(pos + 3 * size_t(idx[3 * size_t(t) + 0]),
                            pos + 3 * size_t(idx[3 * size_t(t) + 1]),
                            pos + 3 * size_t(idx[3 * size_t(t) + 2]), tl)) {
        atomicOr(w.bad, 1u);
        return;
    }
    Light_info li;
    for (int k = 0; k < 3; ++k) {
        w.tri_lo[3 * size_t(j) + k]   = tl.lo[k];
        w.tri_hi[3 * size_t(j) + k]   = tl.hi[k];
        w.tri_cone[4 * size_t(j) + k] = tl.n[k];
        li.center[k]                  = tl.center[k];
        li.cone[k]                    = tl.n[k];
    }
    w.tri_cone[4 * size_t(j) + 3] = 1.f;
    w.tri_power[j]                = tl.area;
    li.radius    = tl.radius;
    li.cone[3]   = 1.f;
    li.power     = tl.area;
    li.two_sided = 0;
    li.triangle  = t;
    li.area      = tl.area;
    plights[j]   = li;
}

// Part::init of every part and its light record in the scene tree.
__global__ void part_lights(float const* pos, uint32_t const* idx, uint32_t const* emit_tris,
                            float const* radiance, uint32_t num_parts, Emitter_work w,
                            Light_info* infos) {
    uint32_t const r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= num_parts) {
        return;
    }
    uint32_t const first = w.part_first[r];
    uint32_t const n     = w.part_first[r + 1] - first;
    lb::Part_light pl;
    lb::part_light(n,
                   [&](uint32_t i) {
                       uint32_t const t = emit_tris[first + i];
                       lb::Tri_light  tl;
                       lb::triangle_light(pos + 3 * size_t(idx[3 * size_t(t) + 0]),
                                          pos + 3 * size_t(idx[3 * size_t(t) + 1]),
                                          pos + 3 * size_t(idx[3 * size_t(t) + 2]), tl);
                       return tl;
                   },
                   pl);
    Light_info L;
    for (int k = 0; k < 3; ++k) {
        w.part_lo[3 * size_t(r) + k] = pl.lo[k];
        w.part_hi[3 * size_t(r) + k] = pl.hi[k];
        L.center[k]                  = 0.5f * (pl.lo[k] + pl.hi[k]);
    }
    float const ext[3] = {pl.hi[0] - pl.lo[0], pl.hi[1] - pl.lo[1], pl.hi[2] - pl.lo[2]};
    L.radius           = 0.5f * sqrtf((ext[0] * ext[0] + ext[1] * ext[1]) + ext[2] * ext[2]);
    for (int k = 0; k < 4; ++k) {
        w.part_cone[4 * size_t(r) + k] = pl.cone[k];
        L.cone[k]                      = pl.cone[k];
    }
    L.power       = pl.total * radiance[r];
    L.two_sided   = 0;
    L.triangle    = r;
    L.area        = pl.total;
    w.part_pow[r] = L.power;
    w.part_two[r] = 0;
    infos[r]      = L;
}

// The part pools' cdf offsets, absolute: part r's distributions start at 2 first_r.
__global__ void absolute_cdf_offsets(Emitter_work w, dev::Part_info const* infos,
                                     uint32_t num_parts, uint32_t* cdf_offset) {
    uint32_t const r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= num_parts) {
        return;
    }
    dev::Part_info const pi = infos[r];
    for (uint32_t i = pi.node_offset; i < pi.node_offset + pi.num_nodes; ++i) {
        uint32_t const rel = w.cdf_rel[i];
        cdf_offset[i]      = dev::Null == rel ? dev::Null : 2 * pi.light_offset + rel;
    }
}


}  // namespace

void Emitter_work::release() {
    (void)hipFree(tri_lo);
    (void)hipFree(tri_hi);
    (void)hipFree(tri_cone);
    (void)hipFree(tri_power);
    (void)hipFree(part_cone);
    (void)hipFree(part_lo);
    (void)hipFree(part_hi);
    (void)hipFree(part_pow);
    (void)hipFree(part_two);
    (void)hipFree(cdf_rel);
    (void)hipFree(num_cdfs);
    (void)hipFree(part_first);
    (void)hipFree(bad);
    *this = Emitter_work();
}

int32_t update_positions(RenderScene& s, float const* positions, uint32_t num_vertices,
                         Light_forest_scratch& scratch, hipStream_t stream, double* light_ms) {
    mi_bvh_settings const settings = s.bvh_settings;
    if (s.instanced) {
        return fail(MI_ERR_ARGS, "mi_scene_update_positions: an instanced scene (move its "
                                 "instances with mi_scene_update_instances)");
    }
    if (num_vertices != s.num_vertices || !positions) {
        return fail(MI_ERR_ARGS, "mi_scene_update_positions: %u vertices for a scene of %u",
                    num_vertices, s.num_vertices);
    }
    uint32_t const N = s.num_emitters;
    uint32_t const P = s.num_lights;
    Emitter_work&  w = s.emitter_work;
    hipError_t     e = hipSuccess;
    if (N > w.capacity_tris || P > w.capacity_parts || !w.bad) {
        w.release();
        size_t const n = std::max<size_t>(1, N), p = std::max<size_t>(1, P);
        e = hipMalloc(&w.tri_lo, 3 * n * sizeof(float));
        if (!e) e = hipMalloc(&w.tri_hi, 3 * n * sizeof(float));
        if (!e) e = hipMalloc(&w.tri_cone, 4 * n * sizeof(float));
        if (!e) e = hipMalloc(&w.tri_power, n * sizeof(float));
        if (!e) e = hipMalloc(&w.part_cone, 4 * p * sizeof(float));
        if (!e) e = hipMalloc(&w.part_lo, 3 * p * sizeof(float));
        if (!e) e = hipMalloc(&w.part_hi, 3 * p * sizeof(float));
        if (!e) e = hipMalloc(&w.part_pow, p * sizeof(float));
        if (!e) e = hipMalloc(&w.part_two, p);
        if (!e) e = hipMalloc(&w.cdf_rel, 2 * n * sizeof(uint32_t));
        if (!e) e = hipMalloc(&w.num_cdfs, p * sizeof(uint32_t));
        if (!e) e = hipMalloc(&w.part_first, (p + 1) * sizeof(uint32_t));
        if (!e) e = hipMalloc(&w.bad, sizeof(uint32_t));
        if (!e && P) {
            e = hipMemcpy(w.part_first, s.part_first.data(), (P + 1) * sizeof(uint32_t),
                          hipMemcpyHostToDevice);
        }
        if (e) {
            w.release();
            return fail(MI_ERR_HIP, "mi_scene_update_positions: %s", hipGetErrorString(e));
        }
        w.capacity_tris  = N;
        w.capacity_parts = P;
    }
    // the new positions, checked before the scene takes them: every
    // emissive triangle must keep an area
    float* fresh = nullptr;
    e            = hipMalloc(&fresh, std::max<size_t>(1, num_vertices) * 3 * sizeof(float));
    if (!e) e = hipMemcpyAsync(fresh, positions, size_t(num_vertices) * 3 * sizeof(float),
                               hipMemcpyDeviceToDevice, stream);
    if (!e) e = hipMemsetAsync(w.bad, 0, sizeof(uint32_t), stream);
    if (e) {
        (void)hipFree(fresh);
        return fail(MI_ERR_HIP, "mi_scene_update_positions: %s", hipGetErrorString(e));
    }
    auto const t0 = std::chrono::steady_clock::now();
    if (N) {
        hipLaunchKernelGGL(emitter_lights, dim3(blocks(N, 256)), dim3(256), 0, stream, fresh,
                           s.indices, s.emit_tris, N, w, s.part_lights);
        uint32_t bad = 0;
        e = hipMemcpyAsync(&bad, w.bad, sizeof(uint32_t), hipMemcpyDeviceToHost, stream);
        if (!e) e = hipStreamSynchronize(stream);
        if (e || bad) {
            (void)hipFree(fresh);
            // the part lights of the old positions are made again below on
            // success only; restore them by recomputing from the scene's own
            hipLaunchKernelGGL(emitter_lights, dim3(blocks(N, 256)), dim3(256), 0, stream,
                               s.positions, s.indices, s.emit_tris, N, w, s.part_lights);
            (void)hipStreamSynchronize(stream);
            return e ? fail(MI_ERR_HIP, "mi_scene_update_positions: %s", hipGetErrorString(e))
                     : fail(MI_ERR_ARGS, "mi_scene_update_positions: an emissive triangle has "
                                         "no area at the new positions");
        }
    }
    (void)hipFree(s.positions);
    s.positions = fresh;

    if (P) {
        hipLaunchKernelGGL(part_lights, dim3(blocks(P, 64)), dim3(64), 0, stream, s.positions,
                           s.indices, s.emit_tris, s.part_radiance, P, w, s.light_infos);
        // every part's Primitive_tree at once
        Light_forest pin;
        pin.set        = {w.tri_lo, w.tri_hi, w.tri_cone, w.tri_power, nullptr};
        pin.first      = s.part_first.data();
        pin.part_cones = w.part_cone;
        pin.num_roots  = P;
        pin.num_lights = N;
        pin.part       = true;
        std::vector<uint32_t> part_nodes(P);
        Light_forest_out pout;
        pout.nodes          = s.part_nodes;
        pout.middles        = s.part_middles;
        pout.mapping        = s.part_mapping;
        pout.orders         = s.part_orders;
        pout.cdf_offset     = w.cdf_rel;
        pout.cdfs           = s.part_cdfs;
        pout.num_cdfs       = w.num_cdfs;
        pout.num_nodes_host = part_nodes.data();
        if (int32_t const st = build_light_forest(pin, scratch, pout, stream); st) {
            return fail(st, "mi_scene_update_positions: part trees");
        }
        for (uint32_t r = 0; r < P; ++r) {
            s.part_info_host[r].num_nodes = part_nodes[r];
        }
        e = hipMemcpyAsync(s.part_infos, s.part_info_host.data(), P * sizeof(dev::Part_info),
                           hipMemcpyHostToDevice, stream);
        if (e) {
            return fail(MI_ERR_HIP, "mi_scene_update_positions: %s", hipGetErrorString(e));
        }
        hipLaunchKernelGGL(absolute_cdf_offsets, dim3(blocks(P, 64)), dim3(64), 0, stream, w,
                           s.part_infos, P, s.part_cdf_offset);
        // the scene tree over the parts
        uint32_t const first[2] = {0, P};
        Light_forest   sin;
        sin.set        = {w.part_lo, w.part_hi, w.part_cone, w.part_pow, w.part_two};
        sin.first      = first;
        sin.num_roots  = 1;
        sin.num_lights = P;
        uint32_t         scene_nodes = 0;
        Light_forest_out sout;
        sout.nodes          = s.light_nodes;
        sout.middles        = s.light_middles;
        sout.mapping        = s.light_mapping;
        sout.orders         = s.light_orders;
        sout.num_nodes_host = &scene_nodes;
        int32_t st = build_light_forest(sin, scratch, sout, stream);
        if (MI_ERR_UNSPLITTABLE == st) {
            // a node only the host builder's median split divides: build there
            std::vector<float>   lo(3 * size_t(P)), hi(3 * size_t(P)), cn(4 * size_t(P)), pw(P);
            std::vector<uint8_t> two(P);
            e = hipMemcpyAsync(lo.data(), w.part_lo, lo.size() * sizeof(float), hipMemcpyDeviceToHost, stream);
            if (!e) e = hipMemcpyAsync(hi.data(), w.part_hi, hi.size() * sizeof(float), hipMemcpyDeviceToHost, stream);
            if (!e) e = hipMemcpyAsync(cn.data(), w.part_cone, cn.size() * sizeof(float), hipMemcpyDeviceToHost, stream);
            if (!e) e = hipMemcpyAsync(pw.data(), w.part_pow, pw.size() * sizeof(float), hipMemcpyDeviceToHost, stream);
            if (!e) e = hipMemcpyAsync(two.data(), w.part_two, two.size(), hipMemcpyDeviceToHost, stream);
            if (!e) e = hipStreamSynchronize(stream);
            HostLightTree t;
            if (!e) {
                build_light_tree(lo.data(), hi.data(), cn.data(), pw.data(), two.data(), P, t);
                scene_nodes = uint32_t(t.nodes.size());
                e = hipMemcpyAsync(s.light_nodes, t.nodes.data(), t.nodes.size() * sizeof(Light_node), hipMemcpyHostToDevice, stream);
            }
            if (!e) e = hipMemcpyAsync(s.light_middles, t.middles.data(), t.middles.size() * sizeof(uint32_t), hipMemcpyHostToDevice, stream);
            if (!e) e = hipMemcpyAsync(s.light_mapping, t.light_mapping.data(), P * sizeof(uint32_t), hipMemcpyHostToDevice, stream);
            if (!e) e = hipMemcpyAsync(s.light_orders, t.light_orders.data(), P * sizeof(uint32_t), hipMemcpyHostToDevice, stream);
            if (!e) e = hipStreamSynchronize(stream);
            st = e ? MI_ERR_HIP : MI_OK;
        }
        if (st) {
            return fail(st, "mi_scene_update_positions: scene light tree");
        }
        s.num_light_nodes = scene_nodes;
    }
    e = hipStreamSynchronize(stream);
    if (e) {
        return fail(MI_ERR_HIP, "mi_scene_update_positions: %s", hipGetErrorString(e));
    }
    if (light_ms) {
        *light_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    // box_lo / box_hi stay the creation's: they map positions to the guide's
    // inputs, which a trained guide keeps across moves
    // the traced geometry: a new tree from the GPU builder
    DeviceBvh bvh;
    if (int32_t const st = build_bvh(s.positions, num_vertices, 3, s.indices, s.num_triangles,
                                     settings, nullptr, bvh);
        st) {
        return st;
    }
    s.bvh.release();
    s.bvh = bvh;
    return MI_OK;
}

int32_t create_scene_instanced(mi_mesh const* meshes, uint32_t num_meshes,
                               uint32_t const* instance_mesh, float const* world_to_object,
                               float const* object_to_world, uint32_t const* instance_materials,
                               uint32_t num_instances, float const* positions,
                               uint32_t num_vertices, uint32_t const* indices,
                               uint32_t num_triangles, uint32_t const* tri_material,
                               mi_material const* materials, uint32_t num_materials,
                               mi_bvh_settings const& settings, RenderScene& out) {
    for (uint32_t k = 0; k < num_meshes; ++k) {
        mi_mesh const& m = meshes[k];
        if (!m.positions || !m.indices || 0 == m.num_vertices || 0 == m.num_triangles) {
            return fail(MI_ERR_ARGS, "mi_scene_create_instanced: mesh %u is empty", k);
        }
        for (uint64_t i = 0; i < 3ull * m.num_triangles; ++i) {
            if (m.indices[i] >= m.num_vertices) {
                return fail(MI_ERR_ARGS, "mi_scene_create_instanced: mesh %u: index %llu is %u, of %u vertices",
                            k, (unsigned long long)i, m.indices[i], m.num_vertices);
            }
        }
    }
    for (uint32_t i = 0; i < num_instances; ++i) {
        if (instance_mesh[i] >= num_meshes) {
            return fail(MI_ERR_ARGS, "mi_scene_create_instanced: instance %u: mesh %u of %u", i,
                        instance_mesh[i], num_meshes);
        }
        if (instance_materials[i] >= num_materials) {
            return fail(MI_ERR_ARGS, "mi_scene_create_instanced: instance %u: material %u of %u", i,
                        instance_materials[i], num_materials);
        }
        if (is_light(materials[instance_materials[i]])) {
            return fail(MI_ERR_ARGS, "mi_scene_create_instanced: instance %u: emissive material %u",
                        i, instance_materials[i]);
        }
    }

    // the world triangles with their light tree, as a scene of their own
    RenderScene s;
    if (int32_t const st = create_scene(positions, num_vertices, indices, num_triangles, tri_material,
                                        materials, num_materials, settings, s);
        st) {
        return st;
    }
    auto failed = [&s](int32_t st) {
        s.release();
        return st;
    };

    // the meshes: vertex data on the device and one tree each
    std::vector<Mesh_geometry> geo(num_meshes);
    s.mesh_trees.resize(num_meshes);
    for (uint32_t k = 0; k < num_meshes; ++k) {
        mi_mesh const& m   = meshes[k];
        float*         pos = nullptr;
        uint32_t*      idx = nullptr;
        hipError_t     e   = hipMalloc(&pos, size_t(m.num_vertices) * 3 * sizeof(float));
        if (!e) s.mesh_buffers.push_back(pos);
        if (!e) e = hipMalloc(&idx, size_t(m.num_triangles) * 3 * sizeof(uint32_t));
        if (!e) s.mesh_buffers.push_back(idx);
        if (!e) e = hipMemcpy(pos, m.positions, size_t(m.num_vertices) * 3 * sizeof(float), hipMemcpyHostToDevice);
        if (!e) e = hipMemcpy(idx, m.indices, size_t(m.num_triangles) * 3 * sizeof(uint32_t), hipMemcpyHostToDevice);
        if (e) {
            return failed(fail(MI_ERR_HIP, "mi_scene_create_instanced: mesh %u: %s", k, hipGetErrorString(e)));
        }
        geo[k] = {pos, idx, nullptr};
        if (int32_t const st = build_bvh(pos, m.num_vertices, 3, idx, m.num_triangles, settings,
                                         nullptr, s.mesh_trees[k]);
            st) {
            return failed(st);
        }
    }

    // the top-level tree over the instances and, last, the world triangles
    uint32_t const                n_all = num_instances + 1;
    std::vector<DeviceBvh const*> trees(num_meshes + 1);
    for (uint32_t k = 0; k < num_meshes; ++k) {
        trees[k] = &s.mesh_trees[k];
    }
    trees[num_meshes] = &s.bvh;
    std::vector<uint32_t> mesh_of(instance_mesh, instance_mesh + num_instances);
    mesh_of.push_back(num_meshes);
    std::vector<float> w2o(world_to_object, world_to_object + 16 * size_t(num_instances));
    std::vector<float> o2w(object_to_world, object_to_world + 16 * size_t(num_instances));
    for (int k = 0; k < 16; ++k) {
        float const id = (k % 5 == 0) ? 1.f : 0.f;
        w2o.push_back(id);
        o2w.push_back(id);
    }
    for (int k = 0; k < 3; ++k) {
        s.world_box_lo[k] = s.box_lo[k];
        s.world_box_hi[k] = s.box_hi[k];
    }
    s.tlas_settings = settings;
    if (int32_t const st = create_tlas(trees.data(), num_meshes + 1, mesh_of.data(), w2o.data(),
                                       o2w.data(), n_all, settings, nullptr, s.tlas);
        st) {
        return failed(st);
    }
    std::vector<uint32_t> mats(instance_materials, instance_materials + num_instances);
    mats.push_back(dev::Null);
    s.mesh_geo_host = geo;
    s.mesh_vertices.resize(num_meshes);
    s.mesh_normals.assign(num_meshes, nullptr);
    for (uint32_t k = 0; k < num_meshes; ++k) {
        s.mesh_vertices[k] = meshes[k].num_vertices;
    }
    hipError_t e = hipMalloc(&s.mesh_geo, std::max<size_t>(1, num_meshes) * sizeof(Mesh_geometry));
    if (!e) e = hipMalloc(&s.inst_o2w, o2w.size() * sizeof(float));
    if (!e) e = hipMalloc(&s.inst_material, mats.size() * sizeof(uint32_t));
    if (!e && num_meshes) e = hipMemcpy(s.mesh_geo, geo.data(), num_meshes * sizeof(Mesh_geometry), hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(s.inst_o2w, o2w.data(), o2w.size() * sizeof(float), hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(s.inst_material, mats.data(), mats.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e) {
        return failed(fail(MI_ERR_HIP, "mi_scene_create_instanced: %s", hipGetErrorString(e)));
    }
    // the guide's position normalization spans every instance
    for (Instance_record const& r : s.tlas.records) {
        for (int k = 0; k < 3; ++k) {
            s.box_lo[k] = std::min(s.box_lo[k], r.lo[k]);
            s.box_hi[k] = std::max(s.box_hi[k], r.hi[k]);
        }
    }
    s.instanced        = true;
    s.emitter_instance = num_instances;
    out                = s;
    return MI_OK;
}

int32_t update_instances(RenderScene& s, float const* world_to_object, float const* object_to_world,
                         uint32_t num_instances) {
    if (!s.instanced) {
        return fail(MI_ERR_ARGS, "mi_scene_update_instances: the scene has no instances");
    }
    uint32_t const n_all = s.tlas.num_instances;
    if (num_instances + 1 != n_all) {
        return fail(MI_ERR_ARGS, "mi_scene_update_instances: %u instances given, the scene has %u",
                    num_instances, n_all - 1);
    }
    uint32_t const                num_meshes = uint32_t(s.mesh_trees.size());
    std::vector<DeviceBvh const*> trees(num_meshes + 1);
    for (uint32_t k = 0; k < num_meshes; ++k) {
        trees[k] = &s.mesh_trees[k];
    }
    trees[num_meshes] = &s.bvh;
    std::vector<uint32_t> mesh_of(n_all);
    for (uint32_t i = 0; i < n_all; ++i) {
        mesh_of[i] = s.tlas.records[i].mesh;
    }
    std::vector<float> w2o(world_to_object, world_to_object + 16 * size_t(num_instances));
    std::vector<float> o2w(object_to_world, object_to_world + 16 * size_t(num_instances));
    for (int k = 0; k < 16; ++k) {
        float const id = (k % 5 == 0) ? 1.f : 0.f;
        w2o.push_back(id);
        o2w.push_back(id);
    }
    // build the new top level beside the old one, which stays in use on failure
    Tlas fresh;
    if (int32_t const st = create_tlas(trees.data(), num_meshes + 1, mesh_of.data(), w2o.data(),
                                       o2w.data(), n_all, s.tlas_settings, nullptr, fresh);
        st) {
        return st;
    }
    if (hipError_t const e = hipMemcpy(s.inst_o2w, o2w.data(), o2w.size() * sizeof(float),
                                       hipMemcpyHostToDevice);
        e) {
        fresh.release();
        return fail(MI_ERR_HIP, "mi_scene_update_instances: %s", hipGetErrorString(e));
    }
    s.tlas.release();
    s.tlas = std::move(fresh);
    // box_lo / box_hi stay the creation's: they map positions to the guide's
    // inputs, which a trained guide keeps across moves
    return MI_OK;
}

void RenderState::release() {
    (void)hipFree(paths);
    (void)hipFree(rays[0]);
    (void)hipFree(rays[1]);
    (void)hipFree(ray_path[0]);
    (void)hipFree(ray_path[1]);
    (void)hipFree(hits);
    (void)hipFree(shadow_rays);
    (void)hipFree(shadows);
    (void)hipFree(occluded);
    (void)hipFree(shadow_range);
    (void)hipFree(medium_event);
    (void)hipFree(medium_p);
    (void)hipFree(nee);
    (void)hipFree(picks);
    (void)hipFree(counters);
    (void)hipFree(guide_x);
    (void)hipFree(guide_raw);
    (void)hipFree(guide_mix);
    (void)hipFree(train_x);
    (void)hipFree(train_dir);
    (void)hipFree(train_w);
    (void)hipFree(shuf_x);
    (void)hipFree(shuf_dir);
    (void)hipFree(shuf_w);
    (void)hipFree(rec_t);
    (void)hipFree(rec_r);
    (void)hipFree(rec_path);
    (void)hipFree(rec_block);
    (void)hipFree(pixel_samples);
    (void)hipFree(sample_rad);
    (void)hipFree(done);
    (void)hipFree(path_spec);
    (void)hipFree(nee_thr4);
    (void)hipFree(shadow_c4);
    if (live_host) {
        (void)hipHostFree(live_host);
    }
    if (graph) {
        (void)hipGraphExecDestroy(graph);
    }
    if (regen_graph) {
        (void)hipGraphExecDestroy(regen_graph);
    }
    if (begin) {
        (void)hipEventDestroy(begin);
    }
    if (end) {
        (void)hipEventDestroy(end);
    }
    if (stream) {
        (void)hipStreamDestroy(stream);
    }
    *this = RenderState();
}

int32_t RenderState::reserve(uint32_t n, uint32_t num_shadow) {
    if (!stream) {
        hipError_t e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
        if (!e) e = hipEventCreateWithFlags(&begin, hipEventDisableTiming);
        if (!e) e = hipEventCreateWithFlags(&end, hipEventDisableTiming);
        if (!e) e = hipMalloc(&counters, Num_counters * sizeof(uint32_t));
        if (e) {
            release();
            return fail(MI_ERR_HIP, "render state: %s", hipGetErrorString(e));
        }
    }
    if ((num_shadow > shadow_capacity || n > capacity) && stream) {
        (void)hipStreamSynchronize(stream);  // earlier waves may still use the buffers
    }
    if (num_shadow > shadow_capacity) {
        (void)hipFree(shadow_rays);
        (void)hipFree(shadows);
        (void)hipFree(occluded);
        (void)hipFree(picks);
        shadow_rays     = nullptr;
        shadows         = nullptr;
        occluded        = nullptr;
        picks           = nullptr;
        shadow_capacity = 0;
        hipError_t e    = hipMalloc(&shadow_rays, size_t(num_shadow) * sizeof(mi_ray));
        if (!e) e = hipMalloc(&shadows, size_t(num_shadow) * sizeof(Shadow));
        if (!e) e = hipMalloc(&occluded, size_t(num_shadow) * sizeof(uint32_t));
        if (!e) e = hipMalloc(&picks, size_t(num_shadow) * sizeof(Pick_slot));
        if (e) {
            return fail(MI_ERR_HIP, "shadow queue of %u rays: %s", num_shadow, hipGetErrorString(e));
        }
        shadow_capacity = num_shadow;
    }
    if (n <= capacity) {
        return MI_OK;
    }
    (void)hipFree(paths);
    (void)hipFree(rays[0]);
    (void)hipFree(rays[1]);
    (void)hipFree(ray_path[0]);
    (void)hipFree(ray_path[1]);
    (void)hipFree(hits);
    (void)hipFree(shadow_range);
    (void)hipFree(medium_event);
    (void)hipFree(medium_p);
    (void)hipFree(nee);
    capacity     = 0;
    hipError_t e = hipMalloc(&paths, size_t(n) * sizeof(Path));
    for (int q = 0; q < 2 && !e; ++q) {
        e = hipMalloc(&rays[q], size_t(n) * sizeof(mi_ray));
        if (!e) e = hipMalloc(&ray_path[q], size_t(n) * sizeof(uint32_t));
    }
    // room for the closest hits of either scene kind
    static_assert(sizeof(mi_instance_hit) >= sizeof(mi_hit));
    if (!e) e = hipMalloc(&hits, size_t(n) * sizeof(mi_instance_hit));
    inst_hits = reinterpret_cast<mi_instance_hit*>(hits);
    if (!e) e = hipMalloc(&shadow_range, size_t(n) * sizeof(uint2));
    if (!e) e = hipMalloc(&medium_event, size_t(n) * sizeof(uint32_t));
    if (!e) e = hipMalloc(&medium_p, size_t(n) * 3 * sizeof(float));
    if (!e) e = hipMalloc(&nee, size_t(n) * sizeof(Nee_vertex));
    if (e) {
        release();
        return fail(MI_ERR_HIP, "render state of %u paths: %s", n, hipGetErrorString(e));
    }
    capacity = n;
    return MI_OK;
}

namespace spectral {

void response_matrix(float minv[9]) {
    // the bands in double, as bands() forms them
    auto band = [](double lambda, double b[3]) {
        double blue = (510.0 - lambda) / 60.0;
        double red  = (lambda - 560.0) / 50.0;
        blue        = blue < 0.0 ? 0.0 : (blue > 1.0 ? 1.0 : blue);
        red         = red < 0.0 ? 0.0 : (red > 1.0 ? 1.0 : red);
        b[0]        = red;
        b[1]        = (1.0 - blue) - red;
        b[2]        = blue;
    };
    double const knots[6] = {380.0, 450.0, 510.0, 560.0, 610.0, 720.0};
    double       m[9]     = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int piece = 0; piece < 5; ++piece) {
        double const a = knots[piece];
        double const c = knots[piece + 1];
        double       ba[3], bm[3], bc[3];
        band(a, ba);
        band(0.5 * (a + c), bm);
        band(c, bc);
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) {
                m[3 * i + j] += (c - a) / 6.0 * (ba[i] * ba[j] + 4.0 * bm[i] * bm[j] + bc[i] * bc[j]);
            }
        }
    }
    for (double& v : m) {
        v /= double(Lambda_range);
    }
    // the inverse by cofactors
    double const c00 = m[4] * m[8] - m[5] * m[7];
    double const c01 = m[5] * m[6] - m[3] * m[8];
    double const c02 = m[3] * m[7] - m[4] * m[6];
    double const det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    double const inv[9] = {c00 / det,
                           (m[2] * m[7] - m[1] * m[8]) / det,
                           (m[1] * m[5] - m[2] * m[4]) / det,
                           c01 / det,
                           (m[0] * m[8] - m[2] * m[6]) / det,
                           (m[2] * m[3] - m[0] * m[5]) / det,
                           c02 / det,
                           (m[1] * m[6] - m[0] * m[7]) / det,
                           (m[0] * m[4] - m[1] * m[3]) / det};
    for (int k = 0; k < 9; ++k) {
        minv[k] = float(inv[k]);
    }
}

}  // namespace spectral

int32_t RenderState::reserve_spectral(uint32_t n, uint32_t num_shadow) {
    if ((n > spectral_capacity || num_shadow > spectral_shadow_capacity) && stream) {
        (void)hipStreamSynchronize(stream);
    }
    if (n > spectral_capacity) {
        (void)hipFree(path_spec);
        (void)hipFree(nee_thr4);
        path_spec         = nullptr;
        nee_thr4          = nullptr;
        spectral_capacity = 0;
        hipError_t e      = hipMalloc(&path_spec, size_t(n) * sizeof(float2));
        if (!e) e = hipMalloc(&nee_thr4, size_t(n) * sizeof(float));
        if (e) {
            return fail(MI_ERR_HIP, "spectral state of %u paths: %s", n, hipGetErrorString(e));
        }
        spectral_capacity = n;
    }
    if (num_shadow > spectral_shadow_capacity) {
        (void)hipFree(shadow_c4);
        shadow_c4                = nullptr;
        spectral_shadow_capacity = 0;
        if (hipError_t const e = hipMalloc(&shadow_c4, size_t(num_shadow) * sizeof(float)); e) {
            return fail(MI_ERR_HIP, "spectral shadow queue of %u rays: %s", num_shadow,
                        hipGetErrorString(e));
        }
        spectral_shadow_capacity = num_shadow;
    }
    return MI_OK;
}

int32_t RenderState::reserve_regeneration(uint64_t samples, uint32_t slots) {
    if (!live_host) {
        if (hipError_t const e = hipHostMalloc(&live_host, sizeof(uint32_t)); e) {
            live_host = nullptr;
            return fail(MI_ERR_HIP, "regeneration flag: %s", hipGetErrorString(e));
        }
    }
    if (samples <= sample_capacity && slots <= done_capacity) {
        return MI_OK;
    }
    if (stream) {
        (void)hipStreamSynchronize(stream);  // earlier batches may still use the buffers
    }
    if (samples > sample_capacity) {
        (void)hipFree(sample_rad);
        sample_rad      = nullptr;
        sample_capacity = 0;
        if (hipError_t const e = hipMalloc(&sample_rad, samples * 3 * sizeof(float)); e) {
            return fail(MI_ERR_HIP, "sample buffer of %llu samples: %s", (unsigned long long)samples,
                        hipGetErrorString(e));
        }
        sample_capacity = samples;
    }
    if (slots > done_capacity) {
        (void)hipFree(done);
        done          = nullptr;
        done_capacity = 0;
        if (hipError_t const e = hipMalloc(&done, size_t(slots) * sizeof(uint32_t)); e) {
            return fail(MI_ERR_HIP, "ended-path list of %u slots: %s", slots, hipGetErrorString(e));
        }
        done_capacity = slots;
    }
    return MI_OK;
}

int32_t RenderState::reserve_pixel_samples(uint64_t entries) {
    if (entries <= pixel_sample_capacity) {
        return MI_OK;
    }
    if (stream) {
        (void)hipStreamSynchronize(stream);  // earlier frames may still read the table
    }
    (void)hipFree(pixel_samples);
    pixel_samples         = nullptr;
    pixel_sample_capacity = 0;
    if (hipError_t const e = hipMalloc(&pixel_samples, entries * sizeof(float2)); e) {
        return fail(MI_ERR_HIP, "pixel sample table (%llu): %s", (unsigned long long)entries,
                    hipGetErrorString(e));
    }
    pixel_sample_capacity = entries;
    return MI_OK;
}

int32_t RenderState::reserve_guiding(uint32_t n, uint32_t records) {
    if ((n > guide_capacity || records > record_capacity) && stream) {
        (void)hipStreamSynchronize(stream);  // earlier waves may still use the buffers
    }
    if (n > guide_capacity) {
        (void)hipFree(guide_x);
        (void)hipFree(guide_raw);
        (void)hipFree(guide_mix);
        guide_x = guide_raw = guide_mix = nullptr;
        guide_capacity      = 0;
        hipError_t e        = hipMalloc(&guide_x, size_t(n) * 9 * sizeof(float));
        if (!e) e = hipMalloc(&guide_raw, size_t(n) * 32 * sizeof(float));
        if (!e) e = hipMalloc(&guide_mix, size_t(n) * dev::Vmf_mixture::Stored * sizeof(float));
        if (e) {
            return fail(MI_ERR_HIP, "guiding buffers of %u paths: %s", n, hipGetErrorString(e));
        }
        guide_capacity = n;
    }
    if (records > record_capacity) {
        (void)hipFree(train_x);
        (void)hipFree(train_dir);
        (void)hipFree(train_w);
        (void)hipFree(rec_t);
        (void)hipFree(rec_r);
        (void)hipFree(rec_path);
        (void)hipFree(rec_block);
        (void)hipFree(shuf_x);
        (void)hipFree(shuf_dir);
        (void)hipFree(shuf_w);
        train_x = train_dir = train_w = rec_t = rec_r = nullptr;
        shuf_x = shuf_dir = shuf_w = nullptr;
        rec_path        = nullptr;
        rec_block       = nullptr;
        record_capacity = 0;
        hipError_t e    = hipMalloc(&train_x, size_t(records) * 9 * sizeof(float));
        if (!e) e = hipMalloc(&shuf_x, size_t(records) * 9 * sizeof(float));
        if (!e) e = hipMalloc(&shuf_dir, size_t(records) * 3 * sizeof(float));
        if (!e) e = hipMalloc(&shuf_w, size_t(records) * sizeof(float));
        if (!e) e = hipMalloc(&train_dir, size_t(records) * 3 * sizeof(float));
        if (!e) e = hipMalloc(&train_w, size_t(records) * sizeof(float));
        if (!e) e = hipMalloc(&rec_t, size_t(records) * 3 * sizeof(float));
        if (!e) e = hipMalloc(&rec_r, size_t(records) * 3 * sizeof(float));
        if (!e) e = hipMalloc(&rec_path, size_t(records) * sizeof(uint32_t));
        if (!e) e = hipMalloc(&rec_block, size_t((records + 255) / 256) * sizeof(uint32_t));
        if (e) {
            return fail(MI_ERR_HIP, "training records (%u): %s", records, hipGetErrorString(e));
        }
        record_capacity = records;
    }
    return MI_OK;
}

int32_t render(RenderScene const& scene, TraceResources& trace, float const* e_table,
               RenderState& state, mi_camera const& cam, mi_render_settings const& set,
               GuideNet* net, mi_guide_settings const* gs, double* accum, hipStream_t stream,
               mi_render_stats* stats, Sensor_filter const* sensor) {
    bool const guided = nullptr != net;
    if (guided && (!gs || !(gs->fraction >= 0.f) || !(gs->fraction <= 1.f) ||
                   (gs->train && (0 == gs->train_steps || !(gs->learning_rate > 0.f))))) {
        return fail(MI_ERR_ARGS, "mi_render_guided: bad guide settings");
    }
    // the camera grid: the image grown by the filter's reach on every side
    uint32_t const border       = sensor ? uint32_t(sensor->border) : 0u;
    uint32_t const grid_w       = cam.width + 2 * border;
    uint64_t const num_pixels64 = uint64_t(grid_w) * (cam.height + 2 * border);
    uint32_t const spw          = set.samples_per_wave;
    uint64_t const paths64      = num_pixels64 * spw;
    if (0 == num_pixels64 || 0 == spw || 0 == set.samples_per_pixel || paths64 > (1ull << 27) ||
        set.samples_per_pixel % spw || set.max_bounces > 64 || !(cam.fov_deg > 0.f) ||
        !(cam.fov_deg < 180.f)) {
        return fail(MI_ERR_ARGS,
                    "mi_render: bad camera or settings (%ux%u, %u spp in waves of %u, "
                    "max_bounces %u)",
                    cam.width, cam.height, set.samples_per_pixel, spw, set.max_bounces);
    }
    uint32_t const num_pixels = uint32_t(num_pixels64);
    uint32_t const num_paths  = uint32_t(paths64);

    for (float const v : cam.position) {
        if (!std::isfinite(v)) {
            return fail(MI_ERR_ARGS, "mi_render: camera position not finite");
        }
    }
    for (float const v : cam.rotation) {
        if (!std::isfinite(v)) {
            return fail(MI_ERR_ARGS, "mi_render: camera rotation not finite");
        }
    }
    // Perspective::on_update (camera_perspective.cpp:204-226) without lens
    // tilt or shift, in fp32: fov in radians as degrees_to_radians forms it
    float const fov_rad = cam.fov_deg * (3.14159265358979323846f / 180.f);
    float const fw      = float(cam.width);
    float const fh      = float(cam.height);
    float const ratio   = fh / fw;
    float const cz      = 1.f / std::tan(0.5f * fov_rad);
    float const lt[3]   = {-1.f, ratio, cz};
    float const rt[3]   = {1.f, ratio, cz};
    float const lb[3]   = {-1.f, -ratio, cz};

    // Golden_ratio pixel samples: one frame of samples_per_pixel per call
    bool const golden = 0 != (set.flags & MI_RENDER_GOLDEN_RATIO);
    if (golden) {
        uint64_t const entries = num_pixels64 * set.samples_per_pixel;
        if (0 != set.first_sample || entries > (1ull << 30)) {
            return fail(MI_ERR_ARGS,
                        "mi_render: Golden_ratio pixel samples need first_sample 0 and at most "
                        "2^30 samples per frame (%llu)",
                        (unsigned long long)entries);
        }
        if (int32_t const st = state.reserve_pixel_samples(entries); st) {
            return st;
        }
    }

    // a vertex sends one shadow ray per picked light: up to Max_lights where
    // the tree splits
    bool const     split_lights = !(set.flags & MI_RENDER_NO_LIGHT_SPLIT) && scene.num_lights > 1;
    uint64_t const shadow64 =
        uint64_t(num_paths) * (split_lights ? std::min(scene.num_lights, dev::Max_lights) : 1u);
    if (shadow64 > (1ull << 30)) {
        return fail(MI_ERR_ARGS, "mi_render: %llu paths per wave need a shadow queue of %llu rays; "
                    "use fewer samples per wave", (unsigned long long)num_paths,
                    (unsigned long long)shadow64);
    }
    uint32_t const num_shadow = uint32_t(shadow64);
    // regenerating waves: plain renders through the box sensor; the frame
    // in equal batches of whole waves whose samples fit the sample buffer
    // (2^27 samples): the largest multiple of spw dividing spp that fits
    bool const regen        = !guided && !sensor && 0 != (set.flags & MI_RENDER_REGENERATE);
    uint32_t   batch_planes = spw;
    if (regen) {
        uint32_t const fit = uint32_t(std::max<uint64_t>(1, (uint64_t(1) << 27) / num_pixels));
        for (uint32_t planes = spw; planes <= std::min(set.samples_per_pixel, fit); planes += spw) {
            if (0 == set.samples_per_pixel % planes) {
                batch_planes = planes;
            }
        }
    }
    uint64_t const batch_samples = uint64_t(batch_planes) * num_pixels;
    if (regen) {
        if (int32_t const st = state.reserve_regeneration(batch_samples, num_paths); st) {
            return st;
        }
    }
    if (int32_t const st = state.reserve(num_paths, num_shadow); st) {
        return st;
    }
    bool const spectral = 0 != (set.flags & MI_RENDER_SPECTRAL);
    if (spectral) {
        if (int32_t const st = state.reserve_spectral(num_paths, num_shadow); st) {
            return st;
        }
    }
    if (int32_t const st = prepare_trace(scene.bvh, trace, std::max(num_paths, num_shadow)); st) {
        return st;
    }
    uint32_t guide_depth  = 0;
    uint32_t record_depth = 0;
    uint32_t records      = 0;
    if (guided) {
        guide_depth  = gs->guide_depth ? std::min(gs->guide_depth, set.max_bounces + 1) : set.max_bounces + 1;
        record_depth = (gs->train && gs->train_steps)
                           ? std::min(std::min(gs->record_depth, set.max_bounces + 1), guide_depth)
                           : 0;
        records      = uint32_t(std::min<uint64_t>(uint64_t(num_paths) * record_depth, 1ull << 26));
        if (int32_t const st = state.reserve_guiding(num_paths, std::max<uint32_t>(records, 1)); st) {
            return st;
        }
    }

    Wave w;
    std::memset(&w, 0, sizeof(w));  // padding too: the bytes key the cached graph
    w.positions    = scene.positions;
    w.normals      = scene.normals;
    w.indices      = scene.indices;
    w.tri_material = scene.tri_material;
    w.materials    = scene.materials;
    w.tree.nodes     = scene.light_nodes;
    w.tree.middles   = scene.light_middles;
    w.tree.mapping   = scene.light_mapping;
    w.tree.orders    = scene.light_orders;
    w.tree.lights    = scene.light_infos;
    w.tree.num_nodes = scene.num_lights ? scene.num_light_nodes : 0;
    w.tri_light      = scene.tri_light;
    w.split_lights   = split_lights ? 1u : 0u;
    w.parts.parts      = scene.part_infos;
    w.parts.nodes      = scene.part_nodes;
    w.parts.middles    = scene.part_middles;
    w.parts.cdf_offset = scene.part_cdf_offset;
    w.parts.cdfs       = scene.part_cdfs;
    w.parts.mapping    = scene.part_mapping;
    w.parts.orders     = scene.part_orders;
    w.parts.lights     = scene.part_lights;
    w.tri_local        = scene.tri_local;
    w.medium         = scene.medium;
    w.medium_event   = state.medium_event;
    w.medium_p       = state.medium_p;
    w.e_table      = e_table;
    for (int k = 0; k < 3; ++k) {
        w.position[k] = cam.position[k];
        w.left_top[k] = lt[k];
        w.d_x[k]      = (rt[k] - lt[k]) / fw;
        w.d_y[k]      = (lb[k] - lt[k]) / fh;
    }
    for (int k = 0; k < 9; ++k) {
        w.rotation[k] = cam.rotation[k];
    }
    w.pixel_samples = golden ? state.pixel_samples : nullptr;
    w.table_spp     = golden ? set.samples_per_pixel : 0u;
    w.width       = cam.width;
    w.height      = cam.height;
    w.grid_w      = grid_w;
    w.border      = int32_t(border);
    if (sensor) {
        w.filter = *sensor;
    }
    w.min_bounces = set.min_bounces;
    w.max_bounces = set.max_bounces;
    w.seed        = set.seed;
    w.num_paths   = num_paths;
    w.batch_samples = regen ? uint32_t(batch_samples) : num_paths;
    w.regenerate    = regen ? 1u : 0u;
    w.sample_rad    = regen ? state.sample_rad : nullptr;
    w.done          = regen ? state.done : nullptr;
    w.num_pixels  = num_pixels;
    w.paths       = static_cast<Path*>(state.paths);
    w.rays[0]     = state.rays[0];
    w.rays[1]     = state.rays[1];
    w.ray_path[0] = state.ray_path[0];
    w.ray_path[1] = state.ray_path[1];
    w.hits        = state.hits;
    w.ihits       = state.inst_hits;
    if (scene.instanced) {
        w.instanced        = 1;
        w.emitter_instance = scene.emitter_instance;
        w.tlas             = tlas_view(scene.tlas);
        w.mesh_geo         = scene.mesh_geo;
        w.inst_o2w         = scene.inst_o2w;
        w.inst_material    = scene.inst_material;
    }
    w.shadow_rays = state.shadow_rays;
    w.shadows     = static_cast<Shadow*>(state.shadows);
    w.occluded    = state.occluded;
    w.shadow_range = state.shadow_range;
    w.nee          = static_cast<Nee_vertex*>(state.nee);
    w.picks        = static_cast<Pick_slot*>(state.picks);
    w.guide_mix    = state.guide_mix;
    w.counters    = state.counters;
    w.fraction        = guided ? gs->fraction : 0.f;
    w.guide_raw       = state.guide_raw;
    w.guide_x         = state.guide_x;
    w.guide_depth     = guide_depth;
    w.guide_medium    = guided && 0 != gs->medium ? 1u : 0u;
    if (spectral) {
        w.spectral  = 1;
        spectral::response_matrix(w.spec_minv);
        w.path_spec = state.path_spec;
        w.nee_thr4  = state.nee_thr4;
        w.shadow_c4 = state.shadow_c4;
    }
    w.record_depth    = record_depth;
    w.record_capacity = records;
    w.train_x         = state.train_x;
    w.train_dir       = state.train_dir;
    w.train_w         = state.train_w;
    w.rec_t           = state.rec_t;
    w.rec_r           = state.rec_r;
    w.rec_path        = state.rec_path;
    w.rec_block       = state.rec_block;
    for (int k = 0; k < 3; ++k) {
        float const ext = scene.box_hi[k] - scene.box_lo[k];
        w.box_lo[k]     = scene.box_lo[k];
        w.box_scale[k]  = ext > 0.f ? 1.f / ext : 0.f;
    }

    hipStream_t const s         = state.stream;
    uint32_t const    waves     = set.samples_per_pixel / spw;
    uint32_t const    grid_path = std::min<uint32_t>(blocks(num_paths, 256), trace.num_cus * 8);
    uint32_t const    grid_shadow = std::min<uint32_t>(blocks(num_shadow, 256), trace.num_cus * 8);
    auto const        start     = std::chrono::steady_clock::now();

    // traversals with their ray counts on the device: the scene's tree, or
    // the top-level tree over instances
    auto trace_closest = [&](mi_ray const* rays, uint32_t const* count) -> int32_t {
        if (scene.instanced) {
            return trace_instances_launch(scene.tlas, trace.num_cus, rays, num_paths, count,
                                          state.inst_hits, false, s);
        }
        return launch_trace_device_count(scene.bvh, trace, rays, count, num_paths, state.hits,
                                         false, s);
    };
    auto trace_any = [&]() -> int32_t {
        if (scene.instanced) {
            return trace_instances_launch(scene.tlas, trace.num_cus, state.shadow_rays, num_shadow,
                                          &state.counters[C_shadow], state.occluded, true, s);
        }
        return launch_trace_device_count(scene.bvh, trace, state.shadow_rays,
                                         &state.counters[C_shadow], num_shadow, state.occluded,
                                         true, s);
    };

    // one wave: every kernel reads its counts and sample base on the device
    bool const training = guided && gs->train && gs->train_steps > 0;
    auto enqueue_wave = [&]() -> int32_t {
        if (training) {
            hipLaunchKernelGGL(clear_records, dim3(grid_path), dim3(256), 0, s, w);
        }
        hipLaunchKernelGGL(raygen, dim3(blocks(num_paths, 256)), dim3(256), 0, s, w);
        uint32_t q = 0;
        for (uint32_t d = 0; d <= set.max_bounces; ++d) {
            if (int32_t const st = trace_closest(state.rays[q], &state.counters[C_queue0 + Counter_line * q]); st) {
                return st;
            }
            if (scene.medium.present) {
                hipLaunchKernelGGL(spectral ? medium_track_spectral : medium_track,
                                   dim3(grid_path), dim3(256), 0, s, w, q);
            }
            hipLaunchKernelGGL(reset_queues, dim3(1), dim3(1), 0, s, state.counters, q ^ 1);
            if (guided && d < guide_depth) {
                hipLaunchKernelGGL(guide_inputs, dim3(grid_path), dim3(256), 0, s, w, q);
                if (int32_t const st = guide_forward_device_count(
                        *net, state.guide_x, &state.counters[C_queue0 + Counter_line * q], num_paths,
                        state.guide_raw, s);
                    st) {
                    return st;
                }
                hipLaunchKernelGGL(spectral ? shade_guided_spectral : shade_guided,
                                   dim3(grid_path), dim3(256), 0, s, w, q);
            } else if (guided && d <= record_depth) {
                // BSDF sampling; closes the records of the vertex before
                hipLaunchKernelGGL(spectral ? shade_guided_spectral : shade_guided,
                                   dim3(grid_path), dim3(256), 0, s, w, q);
            } else {
                hipLaunchKernelGGL(spectral ? shade_spectral : shade, dim3(grid_path), dim3(256),
                                   0, s, w, q);
            }
            hipLaunchKernelGGL(spectral ? nee_spectral : nee, dim3(grid_shadow), dim3(256), 0, s, w);
            hipLaunchKernelGGL(tally, dim3(1), dim3(1), 0, s, state.counters, q);
            if (int32_t const st = trace_any(); st) {
                return st;
            }
            hipLaunchKernelGGL(spectral ? shadow_resolve_spectral : shadow_resolve,
                               dim3(grid_path), dim3(256), 0, s, w, q);
            q ^= 1;
        }
        if (sensor) {
            hipLaunchKernelGGL(resolve_filtered, dim3(blocks(uint64_t(cam.width) * cam.height, 256)),
                               dim3(256), 0, s, w, spw, accum);
        } else {
            hipLaunchKernelGGL(resolve, dim3(blocks(num_pixels, 256)), dim3(256), 0, s, w, spw, accum);
        }
        hipLaunchKernelGGL(advance_base, dim3(1), dim3(1), 0, s, state.counters, spw);
        return MI_OK;
    };

    // online training on the wave's records, all on the device (capturable):
    // training weights, a shuffled copy, train_steps Adam steps over slices
    auto enqueue_training = [&]() -> int32_t {
        uint32_t const record_blocks = (records + 255) / 256;
        hipLaunchKernelGGL(finish_records, dim3(grid_path), dim3(256), 0, s, w);
        hipLaunchKernelGGL(count_records, dim3(record_blocks), dim3(256), 0, s, w);
        hipLaunchKernelGGL(scan_record_blocks, dim3(1), dim3(1024), 0, s, w, record_blocks);
        hipLaunchKernelGGL(shuffle_records, dim3(record_blocks), dim3(256), 0, s, w,
                           static_cast<uint32_t const*>(net->step_dev), state.shuf_x, state.shuf_dir,
                           state.shuf_w);
        if (int32_t const st = guide_train_records(
                *net, state.shuf_x, state.shuf_dir, state.shuf_w, &state.counters[C_records], records,
                gs->train_steps, gs->batch, gs->learning_rate,
                reinterpret_cast<unsigned long long*>(state.counters + C_totals + 4), s);
            st) {
            return st;
        }
        hipLaunchKernelGGL(reset_records, dim3(1), dim3(1), 0, s, state.counters);
        return MI_OK;
    };

    // one bounce of a regenerating batch: trace queue q, shade into q ^ 1,
    // resolve the shadow rays, restart the ended paths' slots into q ^ 1
    auto enqueue_bounce = [&](uint32_t q) -> int32_t {
        if (int32_t const st = trace_closest(state.rays[q], &state.counters[C_queue0 + Counter_line * q]); st) {
            return st;
        }
        if (scene.medium.present) {
            hipLaunchKernelGGL(spectral ? medium_track_spectral : medium_track, dim3(grid_path),
                               dim3(256), 0, s, w, q);
        }
        hipLaunchKernelGGL(reset_queues, dim3(1), dim3(1), 0, s, state.counters, q ^ 1);
        hipLaunchKernelGGL(spectral ? shade_spectral : shade, dim3(grid_path), dim3(256), 0, s, w, q);
        hipLaunchKernelGGL(spectral ? nee_spectral : nee, dim3(grid_shadow), dim3(256), 0, s, w);
        hipLaunchKernelGGL(tally, dim3(1), dim3(1), 0, s, state.counters, q);
        if (int32_t const st = trace_any(); st) {
            return st;
        }
        hipLaunchKernelGGL(spectral ? shadow_resolve_spectral : shadow_resolve, dim3(grid_path),
                           dim3(256), 0, s, w, q);
        hipLaunchKernelGGL(regenerate, dim3(grid_path), dim3(256), 0, s, w, q ^ 1);
        return MI_OK;
    };

    // order after the caller's earlier work on its stream
    hipError_t e = hipEventRecord(state.begin, stream);
    if (!e) e = hipStreamWaitEvent(s, state.begin, 0);
    if (e) {
        return fail(MI_ERR_HIP, "mi_render: %s", hipGetErrorString(e));
    }
    hipLaunchKernelGGL(set_base, dim3(1), dim3(1), 0, s, state.counters, set.first_sample);
    if (golden) {
        hipLaunchKernelGGL(golden_ratio_samples, dim3(blocks(num_pixels, 256)), dim3(256), 0, s,
                           state.pixel_samples, num_pixels, set.samples_per_pixel);
    }
    if (guided) {
        hipLaunchKernelGGL(reset_records, dim3(1), dim3(1), 0, s, state.counters);
    }

    if (regen) {
        // Bounces run until the queue empties: a slot whose path ends starts
        // the batch's next sample, so the queue stays full until the batch's
        // samples run out.  Two bounces (queues 0 -> 1 -> 0) form one
        // captured graph, replayed until the queue to trace next is empty;
        // every path ends within max_bounces + 1 bounces, which bounds the
        // replays.
        uint32_t const live_slots = uint32_t(std::min<uint64_t>(num_paths, batch_samples));
        uint64_t const max_pairs =
            (uint64_t(set.max_bounces) + 2) * ((batch_samples + live_slots - 1) / live_slots + 1) + 8;
        bool const use_graph = 0 == (set.flags & MI_RENDER_NO_GRAPH);
        if (use_graph) {
            std::vector<uint8_t> key(sizeof(Wave));
            std::memcpy(key.data(), &w, sizeof(Wave));
            auto put = [&key](auto const& v) {
                uint8_t const* b = reinterpret_cast<uint8_t const*>(&v);
                key.insert(key.end(), b, b + sizeof(v));
            };
            put(grid_path);
            put(num_shadow);
            put(scene.bvh.nodes);
            put(scene.bvh.tris);
            put(trace.counter);
            put(trace.spill);
            put(trace.lds_entries);
            put(trace.kernel);
            put(trace.block_size);
            put(trace.blocks_per_cu);
            put(trace.refill_lanes);
            put(trace.claim_batch);
            if (!state.regen_graph || key != state.regen_key) {
                if (state.regen_graph) {
                    (void)hipStreamSynchronize(s);
                    (void)hipGraphExecDestroy(state.regen_graph);
                    state.regen_graph = nullptr;
                    state.regen_key.clear();
                }
                hipGraph_t graph = nullptr;
                e                = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
                if (e) {
                    return fail(MI_ERR_HIP, "mi_render: begin capture: %s", hipGetErrorString(e));
                }
                int32_t st = enqueue_bounce(0);
                if (!st) {
                    st = enqueue_bounce(1);
                }
                hipError_t const ce = hipStreamEndCapture(s, &graph);
                if (st) {
                    if (graph) {
                        (void)hipGraphDestroy(graph);
                    }
                    return st;
                }
                if (ce) {
                    return fail(MI_ERR_HIP, "mi_render: end capture: %s", hipGetErrorString(ce));
                }
                e = hipGraphInstantiate(&state.regen_graph, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                if (e) {
                    state.regen_graph = nullptr;
                    return fail(MI_ERR_HIP, "mi_render: graph instantiate: %s", hipGetErrorString(e));
                }
                state.regen_key = std::move(key);
            }
        }
        for (uint32_t done_planes = 0; done_planes < set.samples_per_pixel; done_planes += batch_planes) {
            hipLaunchKernelGGL(raygen, dim3(blocks(num_paths, 256)), dim3(256), 0, s, w);
            uint64_t pairs = 0;
            for (;;) {
                if (use_graph) {
                    e = hipGraphLaunch(state.regen_graph, s);
                    if (e) {
                        return fail(MI_ERR_HIP, "mi_render: graph launch: %s", hipGetErrorString(e));
                    }
                } else {
                    for (uint32_t q = 0; q < 2; ++q) {
                        if (int32_t const st = enqueue_bounce(q); st) {
                            return st;
                        }
                    }
                }
                e = hipMemcpyAsync(state.live_host, &state.counters[C_queue0], sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, s);
                if (!e) e = hipStreamSynchronize(s);
                if (e) {
                    return fail(MI_ERR_HIP, "mi_render: %s", hipGetErrorString(e));
                }
                if (0 == *state.live_host) {
                    break;
                }
                if (++pairs > max_pairs) {
                    return fail(MI_ERR_HIP, "mi_render: paths still live after %llu bounces",
                                (unsigned long long)(2 * pairs));
                }
            }
            hipLaunchKernelGGL(resolve_samples, dim3(blocks(num_pixels, 256)), dim3(256), 0, s, w,
                               batch_planes, spw, accum);
            hipLaunchKernelGGL(advance_base, dim3(1), dim3(1), 0, s, state.counters, batch_planes);
        }
    } else if (set.flags & MI_RENDER_NO_GRAPH) {
        for (uint32_t wave = 0; wave < waves; ++wave) {
            if (int32_t const st = enqueue_wave(); st) {
                return st;
            }
            if (training) {
                if (int32_t const st = enqueue_training(); st) {
                    return st;
                }
            }
        }
    } else {
        // The wave's graph is kept between calls: everything its kernels take
        // (the wave description, buffers, launch shapes, training settings)
        // forms the key, and an equal key replays the cached graph.
        std::vector<uint8_t> key(sizeof(Wave));
        std::memcpy(key.data(), &w, sizeof(Wave));
        auto put = [&key](auto const& v) {
            uint8_t const* b = reinterpret_cast<uint8_t const*>(&v);
            key.insert(key.end(), b, b + sizeof(v));
        };
        put(accum);
        put(sensor != nullptr);  // which resolve ends the wave
        put(grid_path);
        put(num_shadow);
        put(spw);
        put(scene.bvh.nodes);
        put(scene.bvh.tris);
        put(trace.counter);
        put(trace.spill);
        put(trace.lds_entries);
        put(trace.kernel);
        put(trace.block_size);
        put(trace.blocks_per_cu);
        put(trace.refill_lanes);
        put(trace.claim_batch);
        put(training);
        put(net);
        if (training) {
            put(gs->train_steps);
            put(gs->batch);
            put(gs->learning_rate);
            put(records);
            put(state.shuf_x);
        }
        if (!state.graph || key != state.graph_key) {
            if (state.graph) {
                (void)hipStreamSynchronize(s);  // its earlier launches may still run
                (void)hipGraphExecDestroy(state.graph);
                state.graph = nullptr;
                state.graph_key.clear();
            }
            hipGraph_t graph = nullptr;
            e                = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
            if (e) {
                return fail(MI_ERR_HIP, "mi_render: begin capture: %s", hipGetErrorString(e));
            }
            int32_t st = enqueue_wave();
            if (!st && training) {
                st = enqueue_training();
            }
            hipError_t const ce = hipStreamEndCapture(s, &graph);
            if (st) {
                if (graph) {
                    (void)hipGraphDestroy(graph);
                }
                return st;
            }
            if (ce) {
                return fail(MI_ERR_HIP, "mi_render: end capture: %s", hipGetErrorString(ce));
            }
            e = hipGraphInstantiate(&state.graph, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e) {
                state.graph = nullptr;
                return fail(MI_ERR_HIP, "mi_render: graph instantiate: %s", hipGetErrorString(e));
            }
            state.graph_key = std::move(key);
        }
        for (uint32_t wave = 0; wave < waves && !e; ++wave) {
            e = hipGraphLaunch(state.graph, s);
        }
        if (e) {
            return fail(MI_ERR_HIP, "mi_render: graph launch: %s", hipGetErrorString(e));
        }
    }

    e = hipGetLastError();
    if (!e) e = hipEventRecord(state.end, s);
    if (!e) e = hipStreamWaitEvent(stream, state.end, 0);
    if (e) {
        return fail(MI_ERR_HIP, "mi_render: %s", hipGetErrorString(e));
    }
    if (stats) {
        uint64_t totals[3] = {0, 0, 0};
        e = hipMemcpyAsync(totals, state.counters + C_totals, sizeof(totals), hipMemcpyDeviceToHost,
                           s);
        if (!e) e = hipStreamSynchronize(s);
        if (e) {
            return fail(MI_ERR_HIP, "mi_render: %s", hipGetErrorString(e));
        }
        stats->seconds =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        stats->paths       = double(num_pixels) * set.samples_per_pixel;
        stats->extend_rays = double(totals[0]);
        stats->shadow_rays = double(totals[1]);
        stats->trained     = double(totals[2]);
    }
    return MI_OK;
}

}  // namespace mi
