/*
 * ORACLE (test infrastructure only; never linked into the product).
 * CPU restatement of sprout's MIS path tracer for triangle scenes with
 * Lambert, Substitute and one-sided constant-light materials:
 *
 *   Pathtracer_MIS::integrate         core/rendering/integrator/surface/pathtracer_mis.cpp:95-279
 *   sample_lights / evaluate_light    pathtracer_mis.cpp:281-347
 *   connect_light                     pathtracer_mis.cpp:349-380
 *   power_heuristic, predivided_power_heuristic, russian_roulette
 *                                     core/rendering/integrator/integrator_helper.hpp:32-52
 *   Mesh::sample / Mesh::pdf          core/scene/shape/triangle/triangle_mesh.cpp:462-562
 *   Intersection::offset_p            core/scene/prop/prop_intersection.inl:86-121
 *   lambert::Iso                      core/scene/material/lambert/lambert.inl:15-45
 *   rnd::Generator (PCG32)            base/random/generator.inl:16-49
 *   sample_triangle_uniform           base/math/sampling.inl:39-43
 *
 * Paths draw their random numbers from PCG32 started as start(seed,
 * sample * W * H + y * W + x), in this order: 2 for the pixel jitter; per
 * vertex 1 for Russian roulette when it applies, 3 for the light sample when
 * the scene has lights (1 to pick them, then 2 per picked light), 3 for the
 * BSDF sample (see path() for the additions a medium makes).  Lights (one
 * per emissive triangle) are picked by the light
 * tree (light_tree.c), split below depth 3 when asked (sprout's Adaptive
 * light sampling, pathtracer_mis.cpp:423-426); the pinhole camera jitters
 * uniformly inside the pixel.  Closest and any hits come from the oracle's own SBVH (sbvh.c,
 * trace.c), the Substitute material from bsdf.c.
 *
 * Besides the reference's estimator (ORACLE_PT_MIS) it has two more that
 * converge to the same image and so check its weights: ORACLE_PT_BSDF (no
 * light sampling, every emitter hit counts fully) and ORACLE_PT_NEE (light
 * sampling weighted 1 / light_pdf, emitters hit by BSDF samples count only at
 * the first vertex).
 */

#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define SENTINEL 0xFFFFFFFFu
#define PI_INV_F 0.318309886183790671538f
#define DOT_MIN 0.00001f
#define RAY_MAX_T 3.40282347e+38f

enum { KIND_SUBSTITUTE = 0, KIND_LAMBERT = 1, KIND_LIGHT = 2 };
enum { TYPE_DIFFUSE = 5, TYPE_SPECULAR = 17 };

typedef struct {
    uint64_t state;
    uint64_t inc;
} pcg;

static uint32_t pcg_next(pcg* g) {
    uint64_t const old = g->state;
    g->state           = old * 6364136223846793005ull + (g->inc | 1);
    uint32_t const xrs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t const rot = (uint32_t)(old >> 59u);
    return (xrs >> rot) | (xrs << ((0u - rot) & 31u));
}

static void pcg_start(pcg* g, uint64_t state, uint64_t sequence) {
    g->state = 0;
    g->inc   = (sequence << 1u) | 1u;
    pcg_next(g);
    g->state += state;
    pcg_next(g);
}

static float pcg_float(pcg* g) {
    uint32_t bits = (pcg_next(g) & 0x007FFFFFu) | 0x3F800000u;
    float    f;
    memcpy(&f, &bits, 4);
    return f - 1.f;
}

void oracle_pcg_start(oracle_pcg* g, uint64_t seed, uint64_t sequence) {
    pcg_start((pcg*)g, seed, sequence);
}

float oracle_pcg_float(void* g) {
    return pcg_float((pcg*)g);
}

uint32_t oracle_pcg_uint(oracle_pcg* g) {
    return pcg_next((pcg*)g);
}

typedef struct {
    float x, y, z;
} v3;

static v3 mk(float x, float y, float z) {
    v3 r = {x, y, z};
    return r;
}
static v3 add(v3 a, v3 b) {
    return mk(a.x + b.x, a.y + b.y, a.z + b.z);
}
static v3 sub(v3 a, v3 b) {
    return mk(a.x - b.x, a.y - b.y, a.z - b.z);
}
static v3 scale(float s, v3 a) {
    return mk(s * a.x, s * a.y, s * a.z);
}
static float dot(v3 a, v3 b) {
    return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
static v3 cross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static v3 normalize(v3 v) {
    float const s = 1.f / sqrtf(dot(v, v));
    return scale(s, v);
}
static float clampf(float x, float lo, float hi) {
    return x < lo ? lo : (hi < x ? hi : x);
}

static v3 offset_ray(v3 p, v3 n) {
    float const pa[3] = {p.x, p.y, p.z};
    float const na[3] = {n.x, n.y, n.z};
    float       o[3];
    oracle_offset_ray(pa, na, o);
    return mk(o[0], o[1], o[2]);
}

typedef struct {
    float const*            positions;
    uint32_t const*         indices;
    uint32_t const*         tri_material;
    oracle_material const*  materials;
    oracle_bvh              bvh;
    oracle_light_tree const* tree;
    uint32_t*               tri_light;
    uint32_t*               tri_local;
    oracle_camera const*    cam;
    oracle_pt_settings const* set;
    oracle_perspective      persp;
    double*                 accum;
    uint32_t                next_row;
    pthread_mutex_t         lock;
    int32_t                 status;
    float                   minv[9]; /* spectral: the film's response */
} scene_t;

static v3 vertex(scene_t const* s, uint32_t i) {
    float const* p = s->positions + 3 * (size_t)i;
    return mk(p[0], p[1], p[2]);
}

static void tri_vertices(scene_t const* s, uint32_t t, v3* a, v3* b, v3* c) {
    *a = vertex(s, s->indices[3 * (size_t)t + 0]);
    *b = vertex(s, s->indices[3 * (size_t)t + 1]);
    *c = vertex(s, s->indices[3 * (size_t)t + 2]);
}

/* hit: t, u, v, triangle, instance (SENTINEL without instances) */
static int32_t closest(scene_t const* s, v3 o, v3 d, float max_t, uint32_t hit[5]) {
    float const ray[8] = {o.x, o.y, o.z, 0.f, d.x, d.y, d.z, max_t};
    if (s->set->instances) {
        return oracle_instances_trace(s->set->instances->tlas, ray, 1, 0, hit);
    }
    hit[4] = SENTINEL;
    return oracle_trace(s->bvh.nodes, s->bvh.slots, s->positions, s->indices, ray, 1, 0,
                        ORACLE_RCP_PINNED, hit, NULL, 1);
}

static int32_t occluded(scene_t const* s, v3 o, v3 d, float max_t, uint32_t* occ) {
    float const ray[8] = {o.x, o.y, o.z, 0.f, d.x, d.y, d.z, max_t};
    if (s->set->instances) {
        return oracle_instances_trace(s->set->instances->tlas, ray, 1, 1, occ);
    }
    return oracle_trace(s->bvh.nodes, s->bvh.slots, s->positions, s->indices, ray, 1, 1,
                        ORACLE_RCP_PINNED, occ, NULL, 1);
}

/* The hit of a mesh instance: the triangle interpolated in object space and
 * moved to world space, its normal normalized, transformed and normalized
 * again (Mesh::intersect, triangle_mesh.cpp:379-392; Indexed_data::normal,
 * triangle_bvh_indexed_data.inl:282-293; a normalized transform of the normal
 * in place of the rotation, equal for similarity transforms). */
static void instance_point(oracle_pt_instances const* pi, uint32_t inst, uint32_t tri, float hu,
                           float hv, v3* p, v3* n, v3* ns) {
    uint32_t const     mi   = pi->tlas->mesh_of[inst];
    oracle_mesh const* mesh = &pi->tlas->meshes[mi];
    float const*       pa   = mesh->positions + 3 * (size_t)mesh->indices[3 * (size_t)tri + 0];
    float const*       pb   = mesh->positions + 3 * (size_t)mesh->indices[3 * (size_t)tri + 1];
    float const*       pc   = mesh->positions + 3 * (size_t)mesh->indices[3 * (size_t)tri + 2];
    v3 const           a    = mk(pa[0], pa[1], pa[2]);
    v3 const           b    = mk(pb[0], pb[1], pb[2]);
    v3 const           c    = mk(pc[0], pc[1], pc[2]);
    float const        bw   = (1.f - hu) - hv;
    v3 const           po   = add(add(scale(bw, a), scale(hu, b)), scale(hv, c));
    v3 const           no   = normalize(cross(sub(b, a), sub(c, a)));
    float const*       m    = pi->o2w + 16 * (size_t)inst;
    float const        pi3[3] = {po.x, po.y, po.z};
    float const        ni3[3] = {no.x, no.y, no.z};
    float              pw[3], nw[3];
    oracle_transform_point(m, pi3, pw);
    oracle_transform_vector(m, ni3, nw);
    *p  = mk(pw[0], pw[1], pw[2]);
    *n  = normalize(mk(nw[0], nw[1], nw[2]));
    *ns = *n;
    float const* nn = pi->normals ? pi->normals[mi] : NULL;
    if (nn) {
        /* the mesh's vertex normals interpolated in object space, normalized,
         * moved to world space as the geometric normal is, normalized again */
        uint32_t const* ix = mesh->indices + 3 * (size_t)tri;
        v3 const na = mk(nn[3 * (size_t)ix[0]], nn[3 * (size_t)ix[0] + 1], nn[3 * (size_t)ix[0] + 2]);
        v3 const nb = mk(nn[3 * (size_t)ix[1]], nn[3 * (size_t)ix[1] + 1], nn[3 * (size_t)ix[1] + 2]);
        v3 const nc = mk(nn[3 * (size_t)ix[2]], nn[3 * (size_t)ix[2] + 1], nn[3 * (size_t)ix[2] + 2]);
        v3 const so = normalize(add(add(scale(bw, na), scale(hu, nb)), scale(hv, nc)));
        float const si3[3] = {so.x, so.y, so.z};
        float       sw[3];
        oracle_transform_vector(m, si3, sw);
        *ns = normalize(mk(sw[0], sw[1], sw[2]));
    }
}

#define KIND_PHASE 0x100 /* scattering vertex in the medium */

typedef struct {
    int32_t kind;
    float   g;          /* phase anisotropy */
    float   params[6];  /* color, roughness, metallic, ior */
    float   frame[12];  /* t, b, n, geo_n */
    v3      n;
    v3      t, b;
    v3      wo;
    float   albedo[3];
} surface;

/* ns: the shading normal (the layer's frame), geo_n: the geometric normal
 * (Sample::same_hemisphere, material_sample.inl:141-143) */
static void surface_set(surface* sf, oracle_material const* m, v3 ns, v3 geo_n, v3 wo) {
    float const na[3] = {ns.x, ns.y, ns.z};
    float const ga[3] = {geo_n.x, geo_n.y, geo_n.z};
    float       t[3], b[3];
    oracle_orthonormal_basis(na, t, b);
    sf->kind = (int32_t)m->kind;
    for (int c = 0; c < 3; ++c) {
        sf->params[c] = m->color[c];
        sf->albedo[c] = (1.f - m->metallic) * m->color[c];
    }
    sf->params[3] = m->roughness;
    sf->params[4] = m->metallic;
    sf->params[5] = m->ior;
    for (int k = 0; k < 3; ++k) {
        sf->frame[k]     = t[k];
        sf->frame[3 + k] = b[k];
        sf->frame[6 + k] = na[k];
        sf->frame[9 + k] = ga[k];
    }
    sf->n  = ns;
    sf->t  = mk(t[0], t[1], t[2]);
    sf->b  = mk(b[0], b[1], b[2]);
    sf->wo = wo;
}

/* out: reflection * n.wi (rgb), pdf */
static void surface_evaluate(surface const* sf, v3 wi, float out[4]) {
    if (KIND_PHASE == sf->kind) {
        float const ph = oracle_phase_hg(dot(sf->wo, wi), sf->g);
        out[0] = out[1] = out[2] = out[3] = ph;
        return;
    }
    if (KIND_LAMBERT == sf->kind) {
        float const n_dot_wi = clampf(dot(sf->n, wi), DOT_MIN, 1.f);
        for (int c = 0; c < 3; ++c) {
            out[c] = n_dot_wi * (PI_INV_F * sf->albedo[c]);
        }
        out[3] = n_dot_wi * PI_INV_F;
        return;
    }
    float const wo[3] = {sf->wo.x, sf->wo.y, sf->wo.z};
    float const w[3]  = {wi.x, wi.y, wi.z};
    oracle_substitute_evaluate(1, sf->params, sf->frame, wo, w, out);
}

/* out: wi, reflection (rgb), pdf, type */
static void surface_sample(surface const* sf, float xl, float x0, float x1, float out[8]) {
    if (KIND_PHASE == sf->kind) {
        float const wo[3] = {sf->wo.x, sf->wo.y, sf->wo.z};
        float       pdf;
        oracle_sample_phase(wo, x0, x1, sf->g, out, &pdf);
        out[3] = out[4] = out[5] = pdf;
        out[6]              = pdf;
        uint32_t const type = TYPE_DIFFUSE;
        memcpy(&out[7], &type, 4);
        return;
    }
    if (KIND_LAMBERT == sf->kind) {
        float xy[2];
        oracle_disk_concentric(x0, x1, xy);
        float const z  = sqrtf(fmaxf(0.f, 1.f - xy[0] * xy[0] - xy[1] * xy[1]));
        v3 const    tw = mk(xy[0] * sf->t.x + xy[1] * sf->b.x + z * sf->n.x,
                            xy[0] * sf->t.y + xy[1] * sf->b.y + z * sf->n.y,
                            xy[0] * sf->t.z + xy[1] * sf->b.z + z * sf->n.z);
        v3 const    wi = normalize(tw);
        float const n_dot_wi = clampf(dot(sf->n, wi), DOT_MIN, 1.f);
        out[0] = wi.x;
        out[1] = wi.y;
        out[2] = wi.z;
        for (int c = 0; c < 3; ++c) {
            out[3 + c] = n_dot_wi * (PI_INV_F * sf->albedo[c]);
        }
        out[6]               = n_dot_wi * PI_INV_F;
        uint32_t const type = TYPE_DIFFUSE;
        memcpy(&out[7], &type, 4);
        return;
    }
    float const wo[3] = {sf->wo.x, sf->wo.y, sf->wo.z};
    float const xi[3] = {xl, x0, x1};
    oracle_substitute_sample(1, sf->params, sf->frame, wo, xi, out);
}

static float power_heuristic(float f, float g) {
    float const f2 = f * f;
    return f2 / (f2 + g * g);
}

static float predivided_power_heuristic(float f, float g) {
    return f / (f * f + g * g);
}

/* a shadow ray of the current vertex, resolved after its direction sample */
typedef struct {
    float contrib[4]; /* rgb, or a spectral path's four wavelengths */
    v3    o;
    v3    wi;
    float dist;
} shadow_t;

/* One camera sample; adds its radiance to rgb.
 *
 * With a medium (Tracking_multi, pathtracer_mis.cpp:215-250): the part of
 * every ray before its hit that lies in the medium box is tracked; the
 * tracking weight multiplies the throughput, a scattering point becomes a
 * vertex that samples the phase function and the light tree over the whole
 * sphere, and an abort ends the path.  The unoccluded shadow rays of a vertex
 * are attenuated by the medium after the vertex's direction sample, in the
 * order their lights were picked.  The generator order per vertex is then:
 * tracking of the ray that reached it, [roulette 1], [light select 1, 2 per
 * picked light], direction 3, transmittance of its shadow rays. */
/* One camera sample of grid pixel (gx, gy) of the (W + 2 border) x
 * (H + 2 border) grid, image pixel (gx - border, gy - border); its generator
 * is started at sample * (grid pixels) + gy * (grid width) + gx, as the
 * reference's worker numbers the pixels of its filter-extended tiles
 * (rendering_worker.cpp:93-104).  rgb += radiance; jitter = the pixel offsets.
 *
 * Spectral (spectral.c): the throughput holds the path's four wavelengths,
 * its hero from a hash of (seed, the generator's sequence), so the generator
 * draws as in rgb; colours, emission and the medium's coefficients are
 * upsampled to the wavelengths, and every contribution reaches rgb through
 * the film's response. */
static int32_t path(scene_t const* s, uint32_t gx, uint32_t gy, uint32_t sample,
                    float const* pixel_uv, double rgb[3], float jitter[2]) {
    oracle_pt_settings const* set    = s->set;
    oracle_medium const*      medium = set->medium;
    uint32_t const W = s->cam->width;
    uint32_t const H = s->cam->height;
    uint32_t const Wg = W + 2 * set->border;
    uint32_t const Hg = H + 2 * set->border;
    int32_t const  mode = set->mode;
    int32_t const  x    = (int32_t)gx - (int32_t)set->border;
    int32_t const  y    = (int32_t)gy - (int32_t)set->border;

    pcg            rng;
    uint64_t const sequence = (uint64_t)sample * ((uint64_t)Wg * Hg) + (uint64_t)gy * Wg + gx;
    pcg_start(&rng, set->seed, sequence);
    int32_t const spectral = set->spectral;
    int32_t const nc       = spectral ? 4 : 3;
    float         wb[4][3]; /* spectral: band weights of the four wavelengths */
    float         cca[4], ccs[4];
    if (spectral) {
        float const hero = oracle_spectral_hero(set->seed, sequence);
        for (int k = 0; k < 4; ++k) {
            oracle_spectral_bands(oracle_spectral_wavelength(hero, k), wb[k]);
            if (medium) {
                cca[k] = oracle_spectral_upsample(wb[k], medium->cc_a);
                ccs[k] = oracle_spectral_upsample(wb[k], medium->cc_s);
            }
        }
    }
    float jx = pcg_float(&rng);
    float jy = pcg_float(&rng);
    if (pixel_uv) {
        /* the pixel's Golden_ratio sample; the draws above are spent as with
         * the random sampler */
        jx = pixel_uv[0];
        jy = pixel_uv[1];
    }
    jitter[0] = jx;
    jitter[1] = jy;

    float oo[3], dd[3];
    oracle_perspective_ray(&s->persp, s->cam->position, s->cam->rotation, x, y, jx, jy, oo, dd);
    v3 o = mk(oo[0], oo[1], oo[2]);
    v3 d = mk(dd[0], dd[1], dd[2]);

    float    throughput[4] = {1.f, 1.f, 1.f, 1.f};
    float    radiance[3]   = {0.f, 0.f, 0.f};
    float    bsdf_pdf      = 0.f;
    int      singular      = 1;
    int      primary       = 1; /* State::Primary_ray */
    int      translucent   = 0; /* the previous vertex was in the medium */
    int      nearest       = 0; /* unfiltered density lookups (Filter::Nearest) */
    float    prev_n[3]     = {0.f, 0.f, 0.f};
    shadow_t shadows[ORACLE_MAX_LIGHTS];

    for (uint32_t depth = 0;; ++depth) {
        uint32_t hit[5];
        if (closest(s, o, d, RAY_MAX_T, hit)) {
            return ORACLE_ERR_STACK;
        }
        float t, hu, hv;
        memcpy(&t, &hit[0], 4);
        memcpy(&hu, &hit[1], 4);
        memcpy(&hv, &hit[2], 4);
        uint32_t const tri  = hit[3];
        uint32_t const inst = hit[4];

        int   scatter = 0;
        v3    hit_p   = mk(0.f, 0.f, 0.f);
        v3    geo_n   = mk(0.f, 0.f, 1.f);
        v3    shading_n = geo_n;
        if (medium) {
            float const oa[3] = {o.x, o.y, o.z};
            float const da[3] = {d.x, d.y, d.z};
            float       t0, t1;
            if (oracle_medium_overlap(medium, oa, da, 0.f, SENTINEL == tri ? RAY_MAX_T : t, &t0, &t1)) {
                float         ts = 0.f;
                float         w[4];
                int32_t const e = oracle_medium_track_n(
                    medium, oa, da, t0, t1, depth, nearest, oracle_pcg_float, &rng, nc,
                    spectral ? cca : medium->cc_a, spectral ? ccs : medium->cc_s, &ts, w);
                for (int k = 0; k < nc; ++k) {
                    throughput[k] *= w[k];
                }
                if (2 == e) {
                    break; /* Event::Abort */
                }
                if (1 == e) {
                    scatter = 1;
                    hit_p   = mk(o.x + ts * d.x, o.y + ts * d.y, o.z + ts * d.z);
                }
            }
        }
        if (!scatter && SENTINEL == tri) {
            break;
        }

        v3 const wo = mk(-d.x, -d.y, -d.z);
        surface  sf;
        v3       p_o;

        if (!scatter) {
            oracle_pt_instances const* pi = set->instances;
            oracle_material const*     m;
            uint32_t                   light = SENTINEL;
            if (pi && inst != pi->world) {
                instance_point(pi, inst, tri, hu, hv, &hit_p, &geo_n, &shading_n);
                m = &s->materials[pi->material[inst]];
            } else {
                v3 a, b, c;
                tri_vertices(s, tri, &a, &b, &c);
                float const bw = (1.f - hu) - hv;
                hit_p          = add(add(scale(bw, a), scale(hu, b)), scale(hv, c));
                geo_n          = normalize(cross(sub(b, a), sub(c, a)));
                m              = &s->materials[s->tri_material[tri]];
                light          = s->tri_light[tri];
            }
            if (!(pi && inst != pi->world)) {
                shading_n = geo_n;
            }
            if (set->normals && !(pi && inst != pi->world)) {
                /* the vertex normals interpolated and normalized
                 * (triangle::interpolate_data, triangle_primitive_mt.inl:274-287) */
                float const* nn = set->normals;
                uint32_t const* ix = s->indices + 3 * (size_t)tri;
                v3 const na = mk(nn[3 * (size_t)ix[0]], nn[3 * (size_t)ix[0] + 1], nn[3 * (size_t)ix[0] + 2]);
                v3 const nb = mk(nn[3 * (size_t)ix[1]], nn[3 * (size_t)ix[1] + 1], nn[3 * (size_t)ix[1] + 2]);
                v3 const nc = mk(nn[3 * (size_t)ix[2]], nn[3 * (size_t)ix[2] + 1], nn[3 * (size_t)ix[2] + 2]);
                float const bw = (1.f - hu) - hv;
                shading_n      = normalize(add(add(scale(bw, na), scale(hu, nb)), scale(hv, nc)));
            }

            if (KIND_LIGHT == m->kind) {
                float const cos_l = dot(geo_n, wo);
                if (cos_l > 0.f) {
                    float          weight = 1.f;
                    uint32_t const l      = light;
                    if (0 != depth && !singular && SENTINEL != l) {
                        if (ORACLE_PT_MIS == mode) {
                            int const   split     = set->split_lights && depth < 3;
                            float const origin[3] = {o.x, o.y, o.z};
                            float const tree_pdf =
                                oracle_light_pdf(s->tree, origin, prev_n, translucent, split, l);
                            /* Part::pdf: the part tree's pdf times the relative area */
                            oracle_part const*  part  = &s->tree->parts[l];
                            uint32_t const      local = s->tri_local[tri];
                            float const         ppdf  = oracle_part_pdf(part, origin, prev_n, translucent, local);
                            oracle_light const* L     = &s->tree->lights[l];
                            oracle_light const* T     = &part->tree.lights[local];
                            float const ls_pdf = (t * t) / (cos_l * L->area) * (ppdf * (L->area / T->area));
                            weight             = power_heuristic(bsdf_pdf, ls_pdf * tree_pdf);
                        } else if (ORACLE_PT_NEE == mode) {
                            weight = 0.f;
                        }
                    }
                    if (spectral) {
                        float c4[4], rgb[3];
                        for (int k = 0; k < 4; ++k) {
                            c4[k] = throughput[k] * (weight * oracle_spectral_upsample(wb[k], m->emission));
                        }
                        oracle_spectral_to_rgb(s->minv, (float const(*)[3])wb, c4, rgb);
                        for (int k = 0; k < 3; ++k) {
                            radiance[k] += rgb[k];
                        }
                    } else {
                        for (int k = 0; k < 3; ++k) {
                            radiance[k] += throughput[k] * (weight * m->emission[k]);
                        }
                    }
                }
                break;
            }
            surface_set(&sf, m, shading_n, geo_n, wo);
        } else {
            memset(&sf, 0, sizeof(sf));
            sf.kind = KIND_PHASE;
            sf.g    = medium->g;
            sf.wo   = wo;
        }

        if (0 != depth) {
            if (depth >= set->max_bounces) {
                break;
            }
            if (depth >= set->min_bounces) {
                float q = fmaxf(throughput[0], fmaxf(throughput[1], throughput[2]));
                if (spectral) {
                    q = fmaxf(q, throughput[3]);
                }
                if (pcg_float(&rng) > q) {
                    break;
                }
                for (int k = 0; k < nc; ++k) {
                    throughput[k] /= q;
                }
            }
        }

        if (!scatter && !(dot(geo_n, wo) > 0.f)) {
            break;
        }

        nearest = !(depth <= 1 || primary);

        /* a scattering vertex is its own shadow-ray origin (subsurface offset_p) */
        p_o = scatter ? hit_p : offset_ray(hit_p, geo_n);
        uint32_t num_shadows = 0;
        if (s->tree->num_nodes) {
            int const      split     = set->split_lights && depth < 3;
            float const    select    = pcg_float(&rng);
            float const    po[3]     = {p_o.x, p_o.y, p_o.z};
            float const    gn[3]     = {geo_n.x, geo_n.y, geo_n.z};
            uint32_t       picks[ORACLE_MAX_LIGHTS];
            float          ppdf[ORACLE_MAX_LIGHTS];
            uint32_t const num_picks =
                oracle_light_random(s->tree, po, gn, scatter, select, split, picks, ppdf);
            for (uint32_t k = 0; k < num_picks; ++k) {
                oracle_light const* L  = &s->tree->lights[picks[k]]; /* the part */
                /* Mesh::sample (triangle_mesh.cpp:462-516): the part's tree
                 * picks a triangle, then a point on it */
                oracle_part const*  part = &s->tree->parts[picks[k]];
                float const         rp   = pcg_float(&rng);
                uint32_t            local;
                float               tpdf;
                oracle_part_random(part, po, gn, scatter, rp, &local, &tpdf);
                oracle_light const* T  = &part->tree.lights[local];
                uint32_t const      lt = T->triangle;
                float const         u1 = pcg_float(&rng);
                float const         u2 = pcg_float(&rng);
                if (ORACLE_PT_BSDF == mode) {
                    continue;
                }
                v3 la, lb, lc;
                tri_vertices(s, lt, &la, &lb, &lc);
                float const su = sqrtf(u1);
                float const lu = 1.f - su;
                float const lv = u2 * su;
                float const lw = (1.f - lu) - lv;
                v3 const    lp = add(add(scale(lw, la), scale(lu, lb)), scale(lv, lc));
                v3 const    ln = normalize(cross(sub(lb, la), sub(lc, la)));

                v3 const    axis = sub(offset_ray(lp, ln), p_o);
                float const sl   = dot(axis, axis);
                float const dist = sqrtf(sl);
                v3 const    wi   = mk(axis.x / dist, axis.y / dist, axis.z / dist);
                float const cl   = -dot(ln, wi);
                if (cl >= DOT_MIN) {
                    float f[4];
                    surface_evaluate(&sf, wi, f);
                    float const light_pdf = (sl / (cl * L->area)) * (tpdf * (L->area / T->area)) * ppdf[k];
                    float const weight    = ORACLE_PT_MIS == mode
                                                ? predivided_power_heuristic(light_pdf, f[3])
                                                : 1.f / light_pdf;
                    oracle_material const* lm = &s->materials[s->tri_material[lt]];
                    shadow_t*              sh = &shadows[num_shadows];
                    if (spectral) {
                        for (int c = 0; c < 4; ++c) {
                            float const el = oracle_spectral_upsample(wb[c], lm->emission);
                            float const fl = oracle_spectral_upsample(wb[c], f);
                            sh->contrib[c] = throughput[c] * (weight * (el * fl));
                        }
                    } else {
                        for (int c = 0; c < 3; ++c) {
                            sh->contrib[c] = throughput[c] * (weight * (lm->emission[c] * f[c]));
                        }
                        sh->contrib[3] = 0.f;
                    }
                    if (sh->contrib[0] > 0.f || sh->contrib[1] > 0.f || sh->contrib[2] > 0.f ||
                        sh->contrib[3] > 0.f) {
                        uint32_t occ = 0;
                        if (occluded(s, p_o, wi, dist, &occ)) {
                            return ORACLE_ERR_STACK;
                        }
                        if (!occ) {
                            sh->o    = p_o;
                            sh->wi   = wi;
                            sh->dist = dist;
                            ++num_shadows;
                        }
                    }
                }
            }
        }

        float const xl = pcg_float(&rng);
        float const x0 = pcg_float(&rng);
        float const x1 = pcg_float(&rng);
        float       bs[8];
        surface_sample(&sf, xl, x0, x1, bs);

        /* the vertex's unoccluded shadow rays, after its direction sample */
        for (uint32_t k = 0; k < num_shadows; ++k) {
            shadow_t const* sh    = &shadows[k];
            float           tr[4] = {1.f, 1.f, 1.f, 1.f};
            if (medium) {
                float const oa[3] = {sh->o.x, sh->o.y, sh->o.z};
                float const da[3] = {sh->wi.x, sh->wi.y, sh->wi.z};
                float       t0, t1;
                if (oracle_medium_overlap(medium, oa, da, 0.f, sh->dist, &t0, &t1) &&
                    !oracle_medium_transmittance_n(medium, oa, da, t0, t1, depth, nearest,
                                                   oracle_pcg_float, &rng, nc,
                                                   spectral ? cca : medium->cc_a,
                                                   spectral ? ccs : medium->cc_s, tr)) {
                    continue;
                }
            }
            if (spectral) {
                float c4[4], rgb[3];
                for (int c = 0; c < 4; ++c) {
                    c4[c] = tr[c] * sh->contrib[c];
                }
                oracle_spectral_to_rgb(s->minv, (float const(*)[3])wb, c4, rgb);
                for (int c = 0; c < 3; ++c) {
                    radiance[c] += rgb[c];
                }
            } else {
                for (int c = 0; c < 3; ++c) {
                    radiance[c] += tr[c] * sh->contrib[c];
                }
            }
        }

        if (!(bs[6] > 0.f)) {
            break;
        }
        uint32_t type;
        memcpy(&type, &bs[7], 4);
        singular = TYPE_SPECULAR == type;
        if (!singular) {
            primary = 0;
        }
        translucent = scatter;
        bsdf_pdf    = bs[6];
        if (spectral) {
            for (int k = 0; k < 4; ++k) {
                throughput[k] *= oracle_spectral_upsample(wb[k], bs + 3) / bs[6];
            }
        } else {
            for (int k = 0; k < 3; ++k) {
                throughput[k] *= bs[3 + k] / bs[6];
            }
        }
        v3 const wi = mk(bs[0], bs[1], bs[2]);
        o = scatter ? hit_p
                    : offset_ray(hit_p, dot(geo_n, wi) > 0.f ? geo_n : mk(-geo_n.x, -geo_n.y, -geo_n.z));
        d         = wi;
        prev_n[0] = geo_n.x;
        prev_n[1] = geo_n.y;
        prev_n[2] = geo_n.z;
    }

    for (int k = 0; k < 3; ++k) {
        rgb[k] += (double)radiance[k];
    }
    return 0;
}

static void* render_rows(void* arg) {
    scene_t* s = (scene_t*)arg;
    uint32_t const Wg = s->cam->width + 2 * s->set->border;
    uint32_t const Hg = s->cam->height + 2 * s->set->border;
    float* const   out = s->set->samples;
    float* const   uv  = s->set->golden_ratio
                             ? (float*)malloc(sizeof(float) * 2 * s->set->samples_per_pixel)
                             : NULL;
    for (;;) {
        pthread_mutex_lock(&s->lock);
        uint32_t const y = s->next_row++;
        pthread_mutex_unlock(&s->lock);
        if (y >= Hg) {
            break;
        }
        for (uint32_t x = 0; x < Wg; ++x) {
            double rgb[3] = {0.0, 0.0, 0.0};
            if (uv) {
                /* the grid pixel's generator sequence, iteration 0 */
                oracle_golden_ratio_camera_samples(s->set->samples_per_pixel,
                                                   (uint64_t)y * Wg + x, uv, NULL, NULL);
            }
            for (uint32_t k = 0; k < s->set->samples_per_pixel; ++k) {
                float        jitter[2];
                float const* puv = uv ? uv + 2 * k : NULL;
                if (out) {
                    double one[3] = {0.0, 0.0, 0.0};
                    if (path(s, x, y, s->set->first_sample + k, puv, one, jitter)) {
                        s->status = ORACLE_ERR_STACK;
                    }
                    float* o = out + 5 * (((size_t)k * Hg + y) * Wg + x);
                    o[0]     = (float)one[0];
                    o[1]     = (float)one[1];
                    o[2]     = (float)one[2];
                    o[3]     = jitter[0];
                    o[4]     = jitter[1];
                } else if (path(s, x, y, s->set->first_sample + k, puv, rgb, jitter)) {
                    s->status = ORACLE_ERR_STACK;
                }
            }
            if (!out) {
                double* a = s->accum + 3 * ((size_t)y * Wg + x);
                for (int c = 0; c < 3; ++c) {
                    a[c] += rgb[c];
                }
            }
        }
    }
    free(uv);
    return NULL;
}

int32_t oracle_render(float const* positions, uint32_t num_vertices, uint32_t const* indices,
                      uint32_t num_triangles, uint32_t const* tri_material,
                      oracle_material const* materials, uint32_t num_materials,
                      oracle_light_tree const* tree, oracle_camera const* cam,
                      oracle_pt_settings const* set, double* accum) {
    if (!positions || !indices || !tri_material || !materials || !tree || !cam || !set ||
        (!accum && !set->samples) || (set->border && !set->samples) || set->border > 8 ||
        (set->golden_ratio && set->first_sample) ||
        0 == cam->width || 0 == cam->height || 0 == num_triangles) {
        return ORACLE_ERR_ARGS;
    }
    for (uint32_t t = 0; t < num_triangles; ++t) {
        if (tri_material[t] >= num_materials) {
            return ORACLE_ERR_ARGS;
        }
    }

    scene_t s;
    memset(&s, 0, sizeof(s));
    s.positions    = positions;
    s.indices      = indices;
    s.tri_material = tri_material;
    s.materials    = materials;
    s.cam          = cam;
    s.set          = set;
    s.accum        = accum;

    int32_t const st = oracle_sbvh_build(positions, num_vertices, indices, num_triangles, 16, 64, 4,
                                         &s.bvh);
    if (st) {
        return st;
    }

    /* the light of each emissive triangle */
    s.tree      = tree;
    s.tri_light = (uint32_t*)malloc(sizeof(uint32_t) * num_triangles);
    for (uint32_t t = 0; t < num_triangles; ++t) {
        s.tri_light[t] = SENTINEL;
    }
    s.tri_local = (uint32_t*)malloc(sizeof(uint32_t) * num_triangles);
    for (uint32_t t = 0; t < num_triangles; ++t) {
        s.tri_local[t] = SENTINEL;
    }
    /* scene light l is an emissive part; its triangles are its part tree's lights */
    for (uint32_t l = 0; l < tree->num_lights && tree->parts; ++l) {
        oracle_part const* part = &tree->parts[l];
        for (uint32_t i = 0; i < part->tree.num_lights; ++i) {
            uint32_t const t = part->tree.lights[i].triangle;
            if (t < num_triangles) {
                s.tri_light[t] = l;
                s.tri_local[t] = i;
            }
        }
    }

    oracle_perspective_setup(cam->width, cam->height, cam->fov_deg, &s.persp);
    oracle_spectral_response(s.minv);

    pthread_mutex_init(&s.lock, NULL);
    uint32_t nt = set->num_threads ? set->num_threads : 1;
    if (nt > 256) {
        nt = 256;
    }
    pthread_t threads[256];
    for (uint32_t i = 0; i < nt; ++i) {
        pthread_create(&threads[i], NULL, render_rows, &s);
    }
    for (uint32_t i = 0; i < nt; ++i) {
        pthread_join(threads[i], NULL);
    }
    pthread_mutex_destroy(&s.lock);

    free(s.tri_light);
    free(s.tri_local);
    oracle_bvh_free(&s.bvh);
    return s.status;
}
