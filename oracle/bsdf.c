/*
 * ORACLE (test infrastructure only; never linked into the product).
 *
 * CPU restatement of sprout's Substitute material sample (Frostbite-normalized
 * Disney diffuse + height-correlated GGX with energy compensation):
 *   substitute::Sample::evaluate / sample     core/scene/material/substitute/substitute_sample.cpp:10-45
 *   Base_closure base_evaluate, diffuse_sample, gloss_sample, pure_gloss_*
 *                                             core/scene/material/substitute/substitute_base_sample.inl:17-238
 *   disney::Iso::reflection / reflect / evaluate  core/scene/material/disney/disney.inl:21-77
 *   ggx distribution_isotropic, visibility_and_g1_wo, pdf_visible, Iso::reflection / reflect,
 *   Aniso::sample (VNDF), ilm_ep_conductor, Min_roughness
 *                                             core/scene/material/ggx/ggx.inl:27-236,469-503
 *   E table (directional albedo, 32 x 32)     core/scene/material/ggx/ggx_integrate.cpp:17-45,202-240
 *   fresnel::schlick / schlick_f0             core/scene/material/fresnel/fresnel.inl:10-21
 *   Layer frame helpers, clamp_dot (Dot_min)  core/scene/material/material_sample.inl:25-79,
 *                                             core/scene/material/material_sample_helper.hpp:17-33
 *   sample_disk_concentric / sample_hemisphere_cosine  base/math/sampling.inl:11-32,77-82
 *   orthonormal_basis                         base/math/vector3.inl:523-534
 *   hammersley / radical_inverse_vdC          base/math/sample_distribution.inl:18-52
 *
 * The reference uses SSE estimates for rsqrt (normalize), sqrt (the GGX
 * visibility term) and a polynomial sincos; this restatement uses correctly
 * rounded sqrtf, 1/sqrtf and sinf/cosf, so results agree with the reference
 * to a few float ulps, not bit for bit.  The energy table is regenerated here
 * by the reference's procedure (1024 Hammersley samples per entry, float
 * accumulation of the grid steps) instead of being copied.
 */

#include "oracle.h"

#include <math.h>
#include <string.h>

#define PI_F 3.14159265358979323846f
#define PI_INV_F 0.318309886183790671538f
#define DOT_MIN 0.00001f
#define MIN_ROUGHNESS 0.01314f
#define E_SIZE 32

typedef struct {
    float x, y, z;
} v3;

typedef struct {
    v3 t, b, n;
} olayer;

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

static v3 normalize(v3 v) {
    float const s = 1.f / sqrtf(dot(v, v));
    return scale(s, v);
}

static float clampf(float x, float lo, float hi) {
    return x < lo ? lo : (hi < x ? hi : x);
}

static float clamp_dot(v3 a, v3 b) {
    return clampf(dot(a, b), DOT_MIN, 1.f);
}

static float clamp_abs_dot(v3 a, v3 b) {
    return clampf(fabsf(dot(a, b)), DOT_MIN, 1.f);
}

static v3 tangent_to_world(olayer const* l, v3 v) {
    return mk(v.x * l->t.x + v.y * l->b.x + v.z * l->n.x,
              v.x * l->t.y + v.y * l->b.y + v.z * l->n.y,
              v.x * l->t.z + v.y * l->b.z + v.z * l->n.z);
}

static v3 world_to_tangent(olayer const* l, v3 v) {
    return mk(v.x * l->t.x + v.y * l->t.y + v.z * l->t.z,
              v.x * l->b.x + v.y * l->b.y + v.z * l->b.z,
              v.x * l->n.x + v.y * l->n.y + v.z * l->n.z);
}

static float pow5(float x) {
    float const x2 = x * x;
    return (x2 * x2) * x;
}

static float lerp(float a, float b, float t) {
    float const u = 1.f - t;
    return u * a + t * b;
}

/* ---- sampling helpers ---------------------------------------------------- */

void oracle_disk_concentric(float u, float v, float out[2]) {
    float const sx = 2.f * u - 1.f;
    float const sy = 2.f * v - 1.f;
    if (0.f == sx && 0.f == sy) {
        out[0] = out[1] = 0.f;
        return;
    }
    float r, theta;
    if (fabsf(sx) > fabsf(sy)) {
        r     = sx;
        theta = (PI_F / 4.f) * (sy / sx);
    } else {
        r     = sy;
        theta = (PI_F / 2.f) - (PI_F / 4.f) * (sx / sy);
    }
    out[0] = cosf(theta) * r;
    out[1] = sinf(theta) * r;
}

static v3 hemisphere_cosine(float u, float v) {
    float xy[2];
    oracle_disk_concentric(u, v, xy);
    float const z = sqrtf(fmaxf(0.f, 1.f - xy[0] * xy[0] - xy[1] * xy[1]));
    return mk(xy[0], xy[1], z);
}

void oracle_orthonormal_basis(float const n[3], float t[3], float b[3]) {
    float const sign = copysignf(1.f, n[2]);  // copysign1, base/math/math.hpp:124-128
    float const c    = -1.f / (sign + n[2]);
    float const d    = n[0] * n[1] * c;
    t[0]             = 1.f + sign * n[0] * n[0] * c;
    t[1]             = sign * d;
    t[2]             = -sign * n[0];
    b[0]             = d;
    b[1]             = sign + n[1] * n[1] * c;
    b[2]             = -n[1];
}

static float radical_inverse_vdc(uint32_t bits) {
    bits = (bits << 16) | (bits >> 16);
    bits = ((bits & 0x55555555u) << 1) | ((bits & 0xAAAAAAAAu) >> 1);
    bits = ((bits & 0x33333333u) << 2) | ((bits & 0xCCCCCCCCu) >> 2);
    bits = ((bits & 0x0F0F0F0Fu) << 4) | ((bits & 0xF0F0F0F0u) >> 4);
    bits = ((bits & 0x00FF00FFu) << 8) | ((bits & 0xFF00FF00u) >> 8);
    return (float)bits * 2.3283064365386963e-10f;
}

/* ---- GGX ----------------------------------------------------------------- */

static float distribution_isotropic(float n_dot_h, float a2) {
    float const d = (n_dot_h * n_dot_h) * (a2 - 1.f) + 1.f;
    return a2 / (PI_F * d * d);
}

/* returns visibility (index 0) and G1 denominator term t_wo + n_dot_wo (index 1) */
static void visibility_and_g1_wo(float n_dot_wi, float n_dot_wo, float a2, float out[2]) {
    float const t_wi = sqrtf(a2 + (1.f - a2) * (n_dot_wi * n_dot_wi));
    float const t_wo = sqrtf(a2 + (1.f - a2) * (n_dot_wo * n_dot_wo));
    out[0]           = 0.5f / (n_dot_wi * t_wo + n_dot_wo * t_wi);
    out[1]           = t_wo + n_dot_wo;
}

static float schlick1(float wo_dot_h, float f0) {
    return f0 + pow5(1.f - wo_dot_h) * (1.f - f0);
}

float oracle_schlick_f0(float n0, float n1) {
    float const t = (n0 - n1) / (n0 + n1);
    return t * t;
}

/* Iso::reflection with a Schlick fresnel per channel */
static void ggx_iso_reflection(float n_dot_wi, float n_dot_wo, float wo_dot_h, float n_dot_h,
                               float alpha, float const f0[3], float refl[3], float* pdf) {
    (void)n_dot_wi;
    float const alpha2 = alpha * alpha;
    float const d      = distribution_isotropic(n_dot_h, alpha2);
    float       g[2];
    visibility_and_g1_wo(n_dot_wi, n_dot_wo, alpha2, g);
    for (int c = 0; c < 3; ++c) {
        refl[c] = d * g[0] * schlick1(wo_dot_h, f0[c]);
    }
    *pdf = (0.5f * d) / g[1];
}

/* Aniso::sample (visible normals) with alpha_x == alpha_y == alpha */
static v3 ggx_sample_h(v3 wo, float alpha, float xi0, float xi1, olayer const* l, float* n_dot_h) {
    v3 const lwo = world_to_tangent(l, wo);
    v3 const v   = normalize(mk(alpha * lwo.x, alpha * lwo.y, lwo.z));

    v3 const cross_v_z = normalize(mk(v.y, -v.x, 0.f));
    v3 const t1        = (v.z < 0.9999f) ? cross_v_z : mk(1.f, 0.f, 0.f);
    v3 const t2        = mk(t1.y * v.z, -t1.x * v.z, t1.x * v.y - t1.y * v.x);

    float const a   = 1.f / (1.f + v.z);
    float const r   = sqrtf(xi0);
    float const phi = (xi1 < a) ? (xi1 / a * PI_F) : (PI_F + (xi1 - a) / (1.f - a) * PI_F);

    float const p1 = r * cosf(phi);
    float const p2 = r * sinf(phi) * ((xi1 < a) ? 1.f : v.z);

    v3 m = add(add(scale(p1, t1), scale(p2, t2)),
               scale(sqrtf(fmaxf(1.f - p1 * p1 - p2 * p2, 0.f)), v));

    m        = normalize(mk(alpha * m.x, alpha * m.y, fmaxf(m.z, 0.f)));
    *n_dot_h = clampf(m.z, 0.f, 1.f);
    return tangent_to_world(l, m);
}

typedef struct {
    v3       wi;
    v3       h;
    float    refl[3];
    float    pdf;
    float    h_dot_wi;
    uint32_t type;
} osample;

enum { TYPE_DIFFUSE_REFLECTION = 5, TYPE_GLOSSY_REFLECTION = 9, TYPE_SPECULAR_REFLECTION = 17 };

/* Iso::reflect: returns n_dot_wi */
static float ggx_iso_reflect(v3 wo, float n_dot_wo, float alpha, float const f0[3], float xi0,
                             float xi1, olayer const* l, osample* out) {
    float    n_dot_h;
    v3 const h        = ggx_sample_h(wo, alpha, xi0, xi1, l, &n_dot_h);
    float const wo_dot_h = clamp_dot(wo, h);
    v3 const wi       = normalize(sub(scale(2.f * wo_dot_h, h), wo));
    float const n_dot_wi = clamp_dot(l->n, wi);

    ggx_iso_reflection(n_dot_wi, n_dot_wo, wo_dot_h, n_dot_h, alpha, f0, out->refl, &out->pdf);
    out->wi       = wi;
    out->h        = h;
    out->h_dot_wi = wo_dot_h;
    out->type     = alpha <= MIN_ROUGHNESS * MIN_ROUGHNESS ? TYPE_SPECULAR_REFLECTION
                                                           : TYPE_GLOSSY_REFLECTION;
    return n_dot_wi;
}

/* ---- energy table --------------------------------------------------------- */

static float e_table[E_SIZE * E_SIZE];
static int   e_table_ready = 0;

static float integrate_f_ss(float alpha, float n_dot_wo, uint32_t num_samples) {
    if (alpha < MIN_ROUGHNESS * MIN_ROUGHNESS) {
        return 1.f;
    }
    float const  one[3] = {1.f, 1.f, 1.f};
    olayer const l      = {{1.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, {0.f, 0.f, 1.f}};
    n_dot_wo            = clampf(n_dot_wo, 0.f, 1.f);
    v3 const wo         = mk(sqrtf(1.f - n_dot_wo * n_dot_wo), 0.f, n_dot_wo);

    float accum = 0.f;
    for (uint32_t i = 0; i < num_samples; ++i) {
        float const xi0 = (float)i / (float)num_samples;
        float const xi1 = radical_inverse_vdc(i);
        osample     s;
        float const n_dot_wi = ggx_iso_reflect(wo, n_dot_wo, alpha, one, xi0, xi1, &l, &s);
        accum += (n_dot_wi * s.refl[0]) / s.pdf;
    }
    return accum / (float)num_samples;
}

void oracle_ggx_e_table(float* out) {
    if (!e_table_ready) {
        float const step  = 1.f / (float)(E_SIZE - 1);
        float       alpha = 0.f;
        for (int a = 0; a < E_SIZE; ++a) {
            float n_dot_wo = 0.f;
            for (int i = 0; i < E_SIZE; ++i) {
                e_table[a * E_SIZE + i] = integrate_f_ss(alpha, n_dot_wo, 1024);
                n_dot_wo += step;
            }
            alpha += step;
        }
        e_table_ready = 1;
    }
    if (out) {
        memcpy(out, e_table, sizeof(e_table));
    }
}

/* Interpolated_function_2D_N::operator()(n_dot_wo, alpha) */
static float e_lookup(float x, float y) {
    oracle_ggx_e_table(NULL);
    x               = fminf(x, 1.f);
    y               = fminf(y, 1.f);
    float const ox  = x * (float)(E_SIZE - 1);
    float const oy  = y * (float)(E_SIZE - 1);
    uint32_t const ix = (uint32_t)ox;
    uint32_t const iy = (uint32_t)oy;
    float const tx  = ox - (float)ix;
    float const ty  = oy - (float)iy;
    uint32_t const col1 = ix + 1 < E_SIZE - 1 ? ix + 1 : E_SIZE - 1;
    uint32_t const row0 = iy * E_SIZE;
    uint32_t const row1 = (iy + 1 < E_SIZE - 1 ? iy + 1 : E_SIZE - 1) * E_SIZE;
    float const c00 = e_table[ix + row0];
    float const c10 = e_table[col1 + row0];
    float const c01 = e_table[ix + row1];
    float const c11 = e_table[col1 + row1];
    float const sx = 1.f - tx;
    float const sy = 1.f - ty;
    return sy * (sx * c00 + tx * c10) + ty * (sx * c01 + tx * c11);
}

static void ilm_ep_conductor(float const f0[3], float n_dot_wo, float alpha, float metallic,
                             float out[3]) {
    float const e = e_lookup(n_dot_wo, alpha);
    for (int c = 0; c < 3; ++c) {
        out[c] = 1.f + (metallic / e - 1.f) * f0[c];
    }
}

/* ---- Disney diffuse ------------------------------------------------------- */

static void disney_evaluate(float h_dot_wi, float n_dot_wi, float n_dot_wo, float alpha,
                            float const color[3], float out[3]) {
    float const energy_bias   = lerp(0.f, 0.5f, alpha);
    float const energy_factor = lerp(1.f, 1.f / 1.53f, alpha);
    float const f_d90         = energy_bias + (2.f * alpha) * (h_dot_wi * h_dot_wi);
    float const fmo           = f_d90 - 1.f;
    float const a             = 1.f + fmo * pow5(1.f - n_dot_wi);
    float const b             = 1.f + fmo * pow5(1.f - n_dot_wo);
    for (int c = 0; c < 3; ++c) {
        out[c] = (a * b * energy_factor) * (PI_INV_F * color[c]);
    }
}

static float disney_reflect(v3 wo, float n_dot_wo, olayer const* l, float alpha,
                            float const color[3], float xi0, float xi1, osample* out) {
    v3 const is = hemisphere_cosine(xi0, xi1);
    v3 const wi = normalize(tangent_to_world(l, is));
    v3 const h  = normalize(add(wo, wi));

    float const h_dot_wi = clamp_dot(h, wi);
    float const n_dot_wi = clamp_dot(l->n, wi);

    disney_evaluate(h_dot_wi, n_dot_wi, n_dot_wo, alpha, color, out->refl);
    out->wi       = wi;
    out->h        = h;
    out->pdf      = n_dot_wi * PI_INV_F;
    out->h_dot_wi = h_dot_wi;
    out->type     = TYPE_DIFFUSE_REFLECTION;
    return n_dot_wi;
}

/* ---- Substitute ------------------------------------------------------------ */

typedef struct {
    olayer layer;
    v3     geo_n;
    v3     wo;
    float  alpha;
    float  albedo[3];
    float  f0[3];
    float  metallic;
} osubstitute;

/* params: color.rgb, roughness, metallic, ior; frame: t, b, n, geo_n */
static void substitute_setup(osubstitute* s, float const* params, float const* frame,
                             float const* wo) {
    s->layer.t = mk(frame[0], frame[1], frame[2]);
    s->layer.b = mk(frame[3], frame[4], frame[5]);
    s->layer.n = mk(frame[6], frame[7], frame[8]);
    s->geo_n   = mk(frame[9], frame[10], frame[11]);
    s->wo      = mk(wo[0], wo[1], wo[2]);

    float const r = params[3] > MIN_ROUGHNESS ? params[3] : MIN_ROUGHNESS;
    s->alpha      = r * r;

    float const metallic = params[4];
    float const f0       = oracle_schlick_f0(params[5], 1.f);
    for (int c = 0; c < 3; ++c) {
        s->albedo[c] = (1.f - metallic) * params[c];
        s->f0[c]     = lerp(f0, params[c], metallic);
    }
    s->metallic = metallic;
}

static int same_hemisphere(osubstitute const* s, v3 v) {
    return dot(s->geo_n, v) > 0.f;
}

static void substitute_evaluate(osubstitute const* s, v3 wi, float out[4]) {
    if (!same_hemisphere(s, s->wo)) {
        out[0] = out[1] = out[2] = out[3] = 0.f;
        return;
    }
    v3 const    h        = normalize(add(s->wo, wi));
    float const wo_dot_h = clamp_dot(s->wo, h);
    float const n_dot_wi = clamp_dot(s->layer.n, wi);
    float const n_dot_wo = clamp_abs_dot(s->layer.n, s->wo);
    float const n_dot_h  = clampf(dot(s->layer.n, h), 0.f, 1.f);

    float ggx[3], ggx_pdf, ilm[3];
    ggx_iso_reflection(n_dot_wi, n_dot_wo, wo_dot_h, n_dot_h, s->alpha, s->f0, ggx, &ggx_pdf);
    ilm_ep_conductor(s->f0, n_dot_wo, s->alpha, s->metallic, ilm);

    if (1.f == s->metallic) {
        for (int c = 0; c < 3; ++c) {
            out[c] = n_dot_wi * (ggx[c] * ilm[c]);
        }
        out[3] = ggx_pdf;
        return;
    }

    float d[3];
    disney_evaluate(wo_dot_h, n_dot_wi, n_dot_wo, s->alpha, s->albedo, d);
    for (int c = 0; c < 3; ++c) {
        out[c] = n_dot_wi * (d[c] + ggx[c] * ilm[c]);
    }
    out[3] = 0.5f * (n_dot_wi * PI_INV_F + ggx_pdf);
}

static void substitute_sample(osubstitute const* s, float xi_lobe, float xi0, float xi1,
                              osample* out) {
    memset(out, 0, sizeof(*out));
    if (!same_hemisphere(s, s->wo)) {
        out->pdf = 0.f;
        return;
    }
    float const n_dot_wo = clamp_abs_dot(s->layer.n, s->wo);
    float       ilm[3];
    ilm_ep_conductor(s->f0, n_dot_wo, s->alpha, s->metallic, ilm);

    if (1.f == s->metallic) {
        float const n_dot_wi = ggx_iso_reflect(s->wo, n_dot_wo, s->alpha, s->f0, xi0, xi1,
                                               &s->layer, out);
        for (int c = 0; c < 3; ++c) {
            out->refl[c] *= n_dot_wi * ilm[c];
        }
        return;
    }

    if (xi_lobe < 0.5f) {
        float const n_dot_wi = disney_reflect(s->wo, n_dot_wo, &s->layer, s->alpha, s->albedo,
                                              xi0, xi1, out);
        float const n_dot_h = clampf(dot(s->layer.n, out->h), 0.f, 1.f);
        float       ggx[3], ggx_pdf;
        ggx_iso_reflection(n_dot_wi, n_dot_wo, out->h_dot_wi, n_dot_h, s->alpha, s->f0, ggx,
                           &ggx_pdf);
        for (int c = 0; c < 3; ++c) {
            out->refl[c] = n_dot_wi * (out->refl[c] + ggx[c] * ilm[c]);
        }
        out->pdf = 0.5f * (out->pdf + ggx_pdf);
    } else {
        float const n_dot_wi = ggx_iso_reflect(s->wo, n_dot_wo, s->alpha, s->f0, xi0, xi1,
                                               &s->layer, out);
        float d[3];
        disney_evaluate(out->h_dot_wi, n_dot_wi, n_dot_wo, s->alpha, s->albedo, d);
        for (int c = 0; c < 3; ++c) {
            out->refl[c] = n_dot_wi * (out->refl[c] * ilm[c] + d[c]);
        }
        out->pdf = 0.5f * (out->pdf + n_dot_wi * PI_INV_F);
    }
}

/* batch entry points, same layout as the reference driver */

void oracle_substitute_evaluate(uint32_t n, float const* params, float const* frames,
                                float const* wo, float const* wi, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        osubstitute s;
        substitute_setup(&s, params + 6 * i, frames + 12 * i, wo + 3 * i);
        substitute_evaluate(&s, mk(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]), out + 4 * i);
    }
}

void oracle_substitute_sample(uint32_t n, float const* params, float const* frames,
                              float const* wo, float const* xi, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        osubstitute s;
        substitute_setup(&s, params + 6 * i, frames + 12 * i, wo + 3 * i);
        osample r;
        substitute_sample(&s, xi[3 * i], xi[3 * i + 1], xi[3 * i + 2], &r);
        float* o = out + 8 * i;
        o[0]     = r.wi.x;
        o[1]     = r.wi.y;
        o[2]     = r.wi.z;
        o[3]     = r.refl[0];
        o[4]     = r.refl[1];
        o[5]     = r.refl[2];
        o[6]     = r.pdf;
        uint32_t const type = r.pdf > 0.f ? r.type : 0u;
        memcpy(&o[7], &type, 4);
    }
}
