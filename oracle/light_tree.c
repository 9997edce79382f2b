/*
 * ORACLE (test infrastructure only; never linked into the product).
 * CPU restatement of sprout's light-tree sampling for finite lights:
 *
 *   importance (cone bound)     core/scene/light/light_tree.cpp:27-62
 *   Node::split                 light_tree.cpp:197-224
 *   Node::random_light / pdf    light_tree.cpp:226-303
 *   Tree::random_light          light_tree.cpp:363-450
 *   Tree::pdf                   light_tree.cpp:564-635
 *   distribution_sample_discrete<4>, distribution_pdf<4>
 *                               base/math/distribution_1d.inl:20-28, 88-131
 *
 * The tree itself comes from the caller (any builder's output in the
 * reference's serialized layout).
 */

#include "oracle.h"

#include <math.h>

#define MAX_SPLIT_DEPTH 5u
#define SPLITTING_THRESHOLD 0.00004096f
#define DOT_MIN 0.00001f
#define MIN_NORMAL 1.17549435e-38f

static float dot3(float const* a, float const* b) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

static float clamped_cos_sub(float cos_a, float cos_b, float sin_a, float sin_b) {
    float const angle = cos_a * cos_b + sin_a * sin_b;
    return (cos_a > cos_b) ? 1.f : angle;
}

static float clamped_sin_sub(float cos_a, float cos_b, float sin_a, float sin_b) {
    float const angle = sin_a * cos_b - sin_b * cos_a;
    return (cos_a > cos_b) ? 0.f : angle;
}

float oracle_light_importance(float const p[3], float const n[3], float const center[3],
                              float const cone[4], float radius, float power, int two_sided,
                              int total_sphere) {
    float const axis[3] = {p[0] - center[0], p[1] - center[1], p[2] - center[2]};
    float const il      = 1.f / sqrtf(dot3(axis, axis));
    float const na[3]   = {il * axis[0], il * axis[1], il * axis[2]};
    float const sin_cu  = fminf(il * radius, 1.f);
    float const cos_cone = cone[3];
    float const dna      = dot3(cone, na);
    float const cos_a    = two_sided ? fabsf(dna) : dna;
    float const cos_n    = -dot3(n, na);

    float const cos_cu   = sqrtf(fmaxf(1.f - sin_cu * sin_cu, MIN_NORMAL));
    float const sin_cone = sqrtf(fmaxf(1.f - cos_cone * cos_cone, MIN_NORMAL));
    float const sin_a    = sqrtf(fmaxf(1.f - cos_a * cos_a, MIN_NORMAL));
    float const sin_n    = sqrtf(fmaxf(1.f - cos_n * cos_n, MIN_NORMAL));

    float const ta = clamped_cos_sub(cos_a, cos_cone, sin_a, sin_cone);
    float const tb = clamped_sin_sub(cos_a, cos_cone, sin_a, sin_cone);
    float const tc = clamped_cos_sub(ta, cos_cu, tb, sin_cu);
    float const tn = clamped_cos_sub(cos_n, cos_cu, sin_n, sin_cu);

    float const ra     = total_sphere ? 1.f : tn;
    float const rb     = fmaxf(tc, 0.f);
    float const id_min = fminf(2.f / radius, il);
    float const base   = power * (id_min * id_min);

    return fmaxf(ra * rb * base, DOT_MIN);
}

static float node_weight(oracle_light_node const* nd, float const p[3], float const n[3], int ts) {
    return oracle_light_importance(p, n, nd->center, nd->cone, nd->radius, nd->power,
                                   (int)((nd->bits >> 1) & 1u), ts);
}

static float light_weight(oracle_light const* l, float const p[3], float const n[3], int ts) {
    return oracle_light_importance(p, n, l->center, l->cone, l->radius, l->power,
                                   (int)l->two_sided, ts);
}

static int node_split(oracle_light_node const* nd, float const p[3]) {
    float const r     = nd->radius;
    float const dv[3] = {p[0] - nd->center[0], p[1] - nd->center[1], p[2] - nd->center[2]};
    float const d     = fminf(sqrtf(dot3(dv, dv)), 1.e6f);
    float const a     = fmaxf(d - r, 0.001f);
    float const b     = d + r;
    float const eg    = 1.f / (a * b);
    float const eg2   = eg * eg;
    float const a3    = a * a * a;
    float const b3    = b * b * b;
    float const e2g   = (b3 - a3) / (3.f * (b - a) * a3 * b3);
    float const vg    = e2g - eg2;
    float const ve    = nd->variance;
    float const ee    = nd->power / (float)nd->num_lights;
    float const s2    = fmaxf(ve * vg + ve * eg2 + ee * ee * vg, 0.f);
    float const ns    = 1.f / (1.f + sqrtf(s2));
    return ns <= SPLITTING_THRESHOLD;
}

static int has_children(oracle_light_node const* nd) {
    return 1u == (nd->bits & 1u);
}

static uint32_t children_or_light(oracle_light_node const* nd) {
    return nd->bits >> 2;
}

static void leaf_cdf(oracle_light_tree const* t, oracle_light_node const* nd, float const p[3],
                     float const n[3], int ts, float cdf[5]) {
    float          w[4];
    uint32_t const begin = children_or_light(nd);
    for (uint32_t i = 0; i < 4; ++i) {
        w[i] = i < nd->num_lights ? light_weight(&t->lights[t->mapping[begin + i]], p, n, ts) : 0.f;
    }
    float const ii = 1.f / (((w[0] + w[1]) + w[2]) + w[3]);
    cdf[0]         = 0.f;
    for (int i = 1; i < 4; ++i) {
        cdf[i] = fmaf(w[i - 1], ii, cdf[i - 1]);
    }
    cdf[4] = 1.f;
}

typedef struct {
    float    pdf;
    float    random;
    uint32_t node;
    uint32_t depth;
} entry;

uint32_t oracle_light_random(oracle_light_tree const* t, float const p[3], float const n[3],
                             int32_t ts, float random, int32_t split, uint32_t* lights,
                             float* pdfs) {
    if (0 == t->num_nodes) {
        return 0;
    }
    entry    stack[(1u << (MAX_SPLIT_DEPTH - 1)) + 1];
    uint32_t sp    = 0;
    uint32_t count = 0;
    entry    e     = {1.f, random, 0, split ? 0u : MAX_SPLIT_DEPTH};
    stack[sp++]    = e;

    while (sp > 0) {
        oracle_light_node const* nd       = &t->nodes[e.node];
        int const                do_split = e.depth < MAX_SPLIT_DEPTH && node_split(nd, p);
        if (has_children(nd)) {
            uint32_t const c0 = children_or_light(nd);
            uint32_t const c1 = c0 + 1;
            if (do_split) {
                e.node      = c0;
                ++e.depth;
                entry const push = {e.pdf, e.random, c1, e.depth};
                stack[sp++]      = push;
            } else {
                float       p0 = node_weight(&t->nodes[c0], p, n, ts);
                float       p1 = node_weight(&t->nodes[c1], p, n, ts);
                float const pt = p0 + p1;
                p0 /= pt;
                p1 /= pt;
                if (e.random < p0) {
                    e.node = c0;
                    e.pdf *= p0;
                    e.random /= p0;
                } else {
                    e.node = c1;
                    e.pdf *= p1;
                    e.random = (e.random - p0) / p1;
                }
                e.depth = MAX_SPLIT_DEPTH;
            }
        } else {
            uint32_t const begin = children_or_light(nd);
            if (do_split) {
                for (uint32_t i = 0; i < nd->num_lights && count < ORACLE_MAX_LIGHTS; ++i) {
                    lights[count] = t->mapping[begin + i];
                    pdfs[count++] = e.pdf;
                }
            } else if (1 == nd->num_lights) {
                if (count < ORACLE_MAX_LIGHTS) {
                    lights[count] = t->mapping[begin];
                    pdfs[count++] = e.pdf;
                }
            } else {
                float cdf[5];
                leaf_cdf(t, nd, p, n, ts, cdf);
                uint32_t it = nd->num_lights;
                for (uint32_t i = 0; i < nd->num_lights; ++i) {
                    if (cdf[i] >= e.random) {
                        it = i;
                        break;
                    }
                }
                uint32_t const off = 0 != it ? it - 1 : 0;
                if (count < ORACLE_MAX_LIGHTS) {
                    lights[count] = t->mapping[begin + off];
                    pdfs[count++] = (cdf[off + 1] - cdf[off]) * e.pdf;
                }
            }
            e = stack[--sp];
        }
    }
    return count;
}

float oracle_light_pdf(oracle_light_tree const* t, float const p[3], float const n[3],
                       int32_t ts, int32_t split, uint32_t light) {
    if (0 == t->num_nodes) {
        return 0.f;
    }
    uint32_t const lo  = t->orders[light];
    float          pdf = 1.f;
    for (uint32_t nid = 0, depth = split ? 0u : MAX_SPLIT_DEPTH;; ++depth) {
        oracle_light_node const* nd       = &t->nodes[nid];
        int const                do_split = depth < MAX_SPLIT_DEPTH && node_split(nd, p);
        if (has_children(nd)) {
            uint32_t const c0     = children_or_light(nd);
            uint32_t const middle = t->middles[nid];
            if (do_split) {
                nid = lo < middle ? c0 : c0 + 1;
            } else {
                float const p0 = node_weight(&t->nodes[c0], p, n, ts);
                float const p1 = node_weight(&t->nodes[c0 + 1], p, n, ts);
                float const pt = p0 + p1;
                if (lo < middle) {
                    nid = c0;
                    pdf *= p0 / pt;
                } else {
                    nid = c0 + 1;
                    pdf *= p1 / pt;
                }
                depth = MAX_SPLIT_DEPTH;
            }
        } else {
            if (do_split || 1 == nd->num_lights) {
                return pdf;
            }
            float cdf[5];
            leaf_cdf(t, nd, p, n, ts, cdf);
            uint32_t const idx = lo - children_or_light(nd);
            return pdf * (cdf[idx + 1] - cdf[idx]);
        }
    }
}

/* ---- Primitive_tree (light_tree.cpp:777-862) ---- */

static float leaf_pick_pdf(float const* cdf, uint32_t n, float random, uint32_t* off) {
    uint32_t it = n;
    for (uint32_t i = 0; i < n; ++i) {
        if (cdf[i] >= random) {
            it = i;
            break;
        }
    }
    *off = 0 != it ? it - 1 : 0;
    return cdf[*off + 1] - cdf[*off];
}

void oracle_part_random(oracle_part const* part, float const p[3], float const n[3], int32_t ts,
                        float random, uint32_t* local, float* out_pdf) {
    oracle_light_tree const* t   = &part->tree;
    float                    pdf = 1.f;
    for (uint32_t nid = 0;;) {
        oracle_light_node const* nd = &t->nodes[nid];
        if (has_children(nd)) {
            uint32_t const c0 = children_or_light(nd);
            float          p0 = node_weight(&t->nodes[c0], p, n, ts);
            float          p1 = node_weight(&t->nodes[c0 + 1], p, n, ts);
            float const    pt = p0 + p1;
            p0 /= pt;
            p1 /= pt;
            if (random < p0) {
                nid = c0;
                pdf *= p0;
                random /= p0;
            } else {
                nid = c0 + 1;
                pdf *= p1;
                random = (random - p0) / p1;
            }
            continue;
        }
        uint32_t const begin = children_or_light(nd);
        uint32_t       off   = 0;
        if (nd->num_lights <= 4) {
            if (1 == nd->num_lights) {
                *local   = t->mapping[begin];
                *out_pdf = pdf;
                return;
            }
            float cdf[5];
            leaf_cdf(t, nd, p, n, ts, cdf);
            float const lp = leaf_pick_pdf(cdf, nd->num_lights, random, &off);
            *local         = t->mapping[begin + off];
            *out_pdf       = lp * pdf;
            return;
        }
        float const lp = leaf_pick_pdf(part->cdfs + part->cdf_offset[nid], nd->num_lights, random, &off);
        *local         = t->mapping[begin + off];
        *out_pdf       = lp * pdf;
        return;
    }
}

float oracle_part_pdf(oracle_part const* part, float const p[3], float const n[3], int32_t ts,
                      uint32_t local) {
    oracle_light_tree const* t   = &part->tree;
    uint32_t const           lo  = t->orders[local];
    float                    pdf = 1.f;
    for (uint32_t nid = 0;;) {
        oracle_light_node const* nd     = &t->nodes[nid];
        uint32_t const           middle = t->middles[nid];
        if (middle > 0) {
            uint32_t const c0 = children_or_light(nd);
            float const    p0 = node_weight(&t->nodes[c0], p, n, ts);
            float const    p1 = node_weight(&t->nodes[c0 + 1], p, n, ts);
            float const    pt = p0 + p1;
            if (lo < middle) {
                nid = c0;
                pdf *= p0 / pt;
            } else {
                nid = c0 + 1;
                pdf *= p1 / pt;
            }
            continue;
        }
        uint32_t const idx = lo - children_or_light(nd);
        if (nd->num_lights <= 4) {
            if (1 == nd->num_lights) {
                return pdf;
            }
            float cdf[5];
            leaf_cdf(t, nd, p, n, ts, cdf);
            return pdf * (cdf[idx + 1] - cdf[idx]);
        }
        float const* cdf = part->cdfs + part->cdf_offset[nid];
        return pdf * (cdf[idx + 1] - cdf[idx]);
    }
}

void oracle_part_cone(float const* normals, float const* areas, uint32_t n, float out[4]) {
    float axis[3] = {0.f, 0.f, 0.f};
    float total   = 0.f;
    for (uint32_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            axis[k] += areas[i] * normals[3 * i + k];
        }
        total += areas[i];
    }
    float ax[3] = {axis[0] / total, axis[1] / total, axis[2] / total};
    float const al = sqrtf((ax[0] * ax[0] + ax[1] * ax[1]) + ax[2] * ax[2]);
    for (int k = 0; k < 3; ++k) {
        ax[k] /= al;
    }
    float angle = 0.f;
    for (uint32_t i = 0; i < n; ++i) {
        float const* nn = normals + 3 * i;
        float const  c  = (ax[0] * nn[0] + ax[1] * nn[1]) + ax[2] * nn[2];
        angle           = fmaxf(angle, acosf(fminf(fmaxf(c, -1.f), 1.f)));
    }
    out[0] = ax[0];
    out[1] = ax[1];
    out[2] = ax[2];
    out[3] = cosf(angle);
}
