/*
 * ORACLE (test infrastructure only; never linked into the product).
 * CPU restatement of sprout's heterogeneous-medium machinery for a
 * one-channel density grid filling an axis-aligned box:
 *
 *   Octree_builder::build / split / serialize
 *                           core/scene/material/volumetric/volumetric_octree_builder.cpp:13-296
 *   Gridtree::intersect     core/scene/material/volumetric/volumetric_octree.cpp:53-129
 *   Node encoding           core/scene/material/volumetric/volumetric_octree.hpp:19-58
 *   AABB::intersect_inside  base/math/aabb.inl:184-227 (minps / maxps NaN filtering)
 *   Grid density            core/scene/material/volumetric/volumetric_grid.cpp:35-66
 *   Linear_3D / Nearest_3D, clamp addressing
 *                           core/image/texture/texture_sampler.cpp:124-188, 256-282,
 *                           core/image/texture/address_mode.hpp:8-28,
 *                           core/image/typed_image.cpp:124-160 (gather order)
 *   bilinear, lerp          base/math/math.hpp:131-136
 *   CM(CC), CM::is_empty    core/scene/material/collision_coefficients.inl:24-41
 *   van_de_hulst, similarity_relation_scale
 *                           collision_coefficients.inl:9-11, core/scene/material/material.cpp:108-139
 *   texture_space_ray       core/rendering/integrator/volume/tracking.inl:12-29
 *   Tracking::tracking      core/rendering/integrator/volume/tracking.cpp:328-375
 *   tracking_transmitted, residual_ratio_tracking_transmitted
 *                           tracking.cpp:32-141
 *   Tracking::transmittance tracking.cpp:143-178
 *   Tracking_multi::integrate (the cell walk and the abort test)
 *                           core/rendering/integrator/volume/tracking_multi.cpp:101-162
 *   phase_hg, sample_phase  core/scene/material/material.inl:15-40,
 *                           core/scene/material/volumetric/volumetric_sample.cpp:15-60
 *   offset_f                core/scene/ray_offset.inl:31-34
 *
 * The medium is a box, not a boundary mesh: a ray that starts outside the
 * box enters it at offset_f of the entry distance (what stepping through a
 * boundary surface does, pathtracer_mis.cpp:185-186); one that starts inside
 * is tracked from its own min_t.  Random numbers come from the caller's
 * generator through a callback, in the reference's order.
 */

#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define LOG2_CELL_DIM 5
#define CELL_DIM (1 << LOG2_CELL_DIM)
#define GRID_EMPTY 0x7FFFFFFFu
#define MIN_MT 1e-10f
#define ABORT_EPSILON 7.5e-4f
#define SR_LOW 16u
#define SR_HIGH 64u
#define PI_F 3.14159265358979323846f

/* ---- builder ---- */

typedef struct build_node {
    struct build_node* children; /* 8 or NULL */
    float              data[4];  /* minorant mu_a, mu_s, majorant mu_a, mu_s */
} build_node;

typedef struct {
    float const* density;
    int32_t      d[3];
    int32_t      uniform; /* every cell split to its 8^3-voxel leaves (flat majorants) */
    float        cm[4];
    uint32_t     num_nodes;
    uint32_t     num_data;
    int32_t      failed;
} splitter;

static int data_empty(float const c[4]) {
    return 0.f == c[2] && 0.f == c[3];
}

static float min3(float const v[3]) {
    return fminf(fminf(v[0], v[1]), v[2]);
}
static float max3(float const v[3]) {
    return fmaxf(fmaxf(v[0], v[1]), v[2]);
}

static void split(splitter* sp, build_node* node, int32_t const lo[3], int32_t const hi[3],
                  uint32_t depth) {
    int32_t const w = (CELL_DIM >> (LOG2_CELL_DIM - 3)) + 1;
    int32_t       minb[3], maxb[3];
    for (int k = 0; k < 3; ++k) {
        minb[k] = lo[k] - 1 > 0 ? lo[k] - 1 : 0;
        maxb[k] = hi[k] + 1 < sp->d[k] ? hi[k] + 1 : sp->d[k];
    }
    float min_density = 1.f;
    float max_density = 0.f;
    for (int32_t z = minb[2]; z < maxb[2]; ++z) {
        for (int32_t y = minb[1]; y < maxb[1]; ++y) {
            for (int32_t x = minb[0]; x < maxb[0]; ++x) {
                float const v =
                    sp->density[((size_t)z * (size_t)sp->d[1] + (size_t)y) * (size_t)sp->d[0] + (size_t)x];
                /* std::min(density, min_density), std::max(density, max_density) */
                min_density = min_density < v ? min_density : v;
                max_density = v < max_density ? max_density : v;
            }
        }
    }
    if (min_density > max_density) {
        min_density = 0.f;
        max_density = 0.f;
    }
    float const diff = max_density - min_density;
    if (LOG2_CELL_DIM - 3 == (int32_t)depth || (!sp->uniform && diff < 0.1f) || maxb[0] - minb[0] < w ||
        maxb[1] - minb[1] < w || maxb[2] - minb[2] < w) {
        node->children = NULL;
        /* mt_epsilon is 0: both branches of the reference store these */
        node->data[0] = min_density * sp->cm[0];
        node->data[1] = min_density * sp->cm[1];
        node->data[2] = max_density * sp->cm[2];
        node->data[3] = max_density * sp->cm[3];
        if (!data_empty(node->data)) {
            ++sp->num_data;
        }
        return;
    }
    ++depth;
    int32_t c[3];
    for (int k = 0; k < 3; ++k) {
        c[k] = lo[k] + ((hi[k] - lo[k]) >> 1);
    }
    node->children = (build_node*)calloc(8, sizeof(build_node));
    if (!node->children) {
        sp->failed = 1;
        return;
    }
    /* the reference's eight sub-boxes, spelled out in its order */
    int32_t const boxes[8][6] = {
        {lo[0], lo[1], lo[2], c[0], c[1], c[2]},    {c[0], lo[1], lo[2], hi[0], c[1], c[2]},
        {lo[0], c[1], lo[2], c[0], hi[1], c[2]},    {c[0], c[1], lo[2], hi[0], hi[1], c[2]},
        {lo[0], lo[1], c[2], c[0], c[1], hi[2]},    {c[0], lo[1], c[2], hi[0], c[1], hi[2]},
        {lo[0], c[1], c[2], c[0], hi[1], hi[2]},    {c[0], c[1], c[2], hi[0], hi[1], hi[2]},
    };
    for (int i = 0; i < 8; ++i) {
        split(sp, &node->children[i], boxes[i], boxes[i] + 3, depth);
    }
    sp->num_nodes += 8;
}

static void serialize(build_node const* node, uint32_t current, uint32_t* next, uint32_t* data,
                      oracle_gridtree* out) {
    if (node->children) {
        out->nodes[current] = (*next << 1) | 1u;
        uint32_t const first = *next;
        *next += 8;
        for (uint32_t i = 0; i < 8; ++i) {
            serialize(&node->children[i], first + i, next, data, out);
        }
    } else if (!data_empty(node->data)) {
        out->nodes[current] = *data << 1;
        memcpy(out->data + 4 * (size_t)*data, node->data, sizeof(node->data));
        ++*data;
    } else {
        out->nodes[current] = GRID_EMPTY << 1;
    }
}

static void free_nodes(build_node* node) {
    if (node->children) {
        for (int i = 0; i < 8; ++i) {
            free_nodes(&node->children[i]);
        }
        free(node->children);
    }
}

int32_t oracle_gridtree_build(float const* density, int32_t const dims[3], float const cc_a[3],
                              float const cc_s[3], oracle_gridtree* out) {
    return oracle_majorant_tree_build(density, dims, cc_a, cc_s, 0, out);
}

int32_t oracle_majorant_tree_build(float const* density, int32_t const dims[3], float const cc_a[3],
                                   float const cc_s[3], int32_t uniform, oracle_gridtree* out) {
    memset(out, 0, sizeof(*out));
    if (!density || dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) {
        return ORACLE_ERR_ARGS;
    }
    int32_t nc[3];
    for (int k = 0; k < 3; ++k) {
        nc[k] = dims[k] >> LOG2_CELL_DIM;
        int32_t const rest = dims[k] - (nc[k] << LOG2_CELL_DIM);
        nc[k] += rest < 1 ? rest : 1;
        out->num_cells[k] = nc[k];
        out->dims[k]      = dims[k];
    }
    uint32_t const cells = (uint32_t)(nc[0] * nc[1] * nc[2]);
    build_node*    grid  = (build_node*)calloc(cells, sizeof(build_node));
    if (!grid) {
        return ORACLE_ERR_ALLOC;
    }
    splitter sp;
    memset(&sp, 0, sizeof(sp));
    sp.density = density;
    sp.uniform = uniform;
    memcpy(sp.d, dims, sizeof(sp.d));
    sp.cm[0] = min3(cc_a);
    sp.cm[1] = min3(cc_s);
    sp.cm[2] = max3(cc_a);
    sp.cm[3] = max3(cc_s);
    int32_t const area = nc[0] * nc[1];
    for (uint32_t i = 0; i < cells; ++i) {
        int32_t c[3];
        c[2]            = (int32_t)i / area;
        int32_t const t = c[2] * area;
        c[1]            = ((int32_t)i - t) / nc[0];
        c[0]            = (int32_t)i - (t + c[1] * nc[0]);
        int32_t lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = c[k] << LOG2_CELL_DIM;
            hi[k] = lo[k] + CELL_DIM;
        }
        split(&sp, &grid[i], lo, hi, 0);
    }
    int32_t st = sp.failed ? ORACLE_ERR_ALLOC : 0;
    if (!st) {
        out->num_nodes = cells + sp.num_nodes;
        out->num_data  = sp.num_data;
        out->nodes     = (uint32_t*)calloc(out->num_nodes, sizeof(uint32_t));
        out->data      = (float*)calloc(4 * (size_t)(out->num_data ? out->num_data : 1), sizeof(float));
        if (!out->nodes || !out->data) {
            st = ORACLE_ERR_ALLOC;
        } else {
            uint32_t next = cells;
            uint32_t data = 0;
            for (uint32_t i = 0; i < cells; ++i) {
                serialize(&grid[i], i, &next, &data, out);
            }
        }
    }
    for (uint32_t i = 0; i < cells; ++i) {
        free_nodes(&grid[i]);
    }
    free(grid);
    if (st) {
        oracle_gridtree_free(out);
    }
    return st;
}

void oracle_gridtree_free(oracle_gridtree* t) {
    free(t->nodes);
    free(t->data);
    t->nodes = NULL;
    t->data  = NULL;
}

/* ---- lookups ---- */

static float saturate(float x) {
    return x < 0.f ? 0.f : (x > 1.f ? 1.f : x);
}

static float voxel(oracle_medium const* m, int32_t x, int32_t y, int32_t z) {
    return m->density[((size_t)z * (size_t)m->dims[1] + (size_t)y) * (size_t)m->dims[0] + (size_t)x];
}

float oracle_medium_density(oracle_medium const* m, float const uvw[3], int32_t nearest) {
    if (nearest) {
        int32_t xyz[3];
        for (int k = 0; k < 3; ++k) {
            int32_t const v = (int32_t)(saturate(uvw[k]) * (float)m->dims[k]);
            xyz[k]          = v < m->dims[k] - 1 ? v : m->dims[k] - 1;
        }
        return voxel(m, xyz[0], xyz[1], xyz[2]);
    }
    int32_t x0[3], x1[3];
    float   st[3];
    for (int k = 0; k < 3; ++k) {
        float const   u  = saturate(uvw[k]) * (float)m->dims[k] - 0.5f;
        float const   fu = floorf(u);
        int32_t const x  = (int32_t)fu;
        int32_t const b  = m->dims[k] - 1;
        x0[k]            = x < 0 ? 0 : x;
        x1[k]            = x >= b ? b : x + 1;
        st[k]            = u - fu;
    }
    float const c[8] = {voxel(m, x0[0], x0[1], x0[2]), voxel(m, x1[0], x0[1], x0[2]),
                        voxel(m, x0[0], x1[1], x0[2]), voxel(m, x1[0], x1[1], x0[2]),
                        voxel(m, x0[0], x0[1], x1[2]), voxel(m, x1[0], x0[1], x1[2]),
                        voxel(m, x0[0], x1[1], x1[2]), voxel(m, x1[0], x1[1], x1[2])};
    float const s  = st[0];
    float const t  = st[1];
    float const _s = 1.f - s;
    float const _t = 1.f - t;
    float const c0 = _t * (_s * c[0] + s * c[1]) + t * (_s * c[2] + s * c[3]);
    float const c1 = _t * (_s * c[4] + s * c[5]) + t * (_s * c[6] + s * c[7]);
    float const q  = st[2];
    return (1.f - q) * c0 + q * c1;
}

typedef struct {
    float o[3];
    float d[3];
    float inv[3];
    float min_t;
    float max_t;
} tray;

static void tray_point(tray const* r, float t, float p[3]) {
    for (int k = 0; k < 3; ++k) {
        p[k] = r->o[k] + t * r->d[k];
    }
}

/* reciprocal3: one Newton step on the estimate, fused, r = 1 / x here */
static float reciprocal(float x) {
    float const r = 1.f / x;
    return fmaf(-(r * r), x, r + r);
}

float oracle_offset_f(float t) {
    if (t < 1.f / 32.f) {
        return t + 1.f / 65536.f;
    }
    int32_t bits;
    memcpy(&bits, &t, 4);
    bits += 256;
    float r;
    memcpy(&r, &bits, 4);
    return r;
}

static float minps(float a, float b) {
    return a < b ? a : b;
}
static float maxps(float a, float b) {
    return a > b ? a : b;
}

static int intersect_inside(float const lo[3], float const hi[3], tray const* r, float* hit_t) {
    float max_lane[3], min_lane[3];
    for (int k = 0; k < 3; ++k) {
        float const l1 = (lo[k] - r->o[k]) * r->inv[k];
        float const l2 = (hi[k] - r->o[k]) * r->inv[k];
        float const a1 = minps(l1, INFINITY);
        float const a2 = minps(l2, INFINITY);
        float const b1 = maxps(l1, -INFINITY);
        float const b2 = maxps(l2, -INFINITY);
        max_lane[k]    = maxps(a1, a2);
        min_lane[k]    = minps(b1, b2);
    }
    float const max_out = minps(minps(max_lane[0], max_lane[1]), max_lane[2]);
    float const min_out = maxps(maxps(min_lane[0], min_lane[1]), min_lane[2]);
    *hit_t              = min_out < r->min_t ? max_out : min_out;
    return max_out >= r->min_t && r->min_t >= min_out && r->max_t >= min_out && max_out >= min_out;
}

/* Gridtree::intersect; returns the leaf's coefficients in cm */
static int gridtree_intersect(oracle_medium const* m, tray* r, float cm[4]) {
    oracle_gridtree const* g = m->tree;
    float                  p[3];
    tray_point(r, r->min_t, p);
    int32_t c[3], v[3];
    for (int k = 0; k < 3; ++k) {
        c[k] = (int32_t)((float)m->dims[k] * p[k]);
        v[k] = c[k] >> LOG2_CELL_DIM;
        if ((uint32_t)v[k] >= (uint32_t)g->num_cells[k]) {
            return 0;
        }
    }
    uint32_t index = ((uint32_t)v[2] * (uint32_t)g->num_cells[1] + (uint32_t)v[1]) *
                         (uint32_t)g->num_cells[0] +
                     (uint32_t)v[0];
    int32_t b0[3], b1[3];
    for (int k = 0; k < 3; ++k) {
        b0[k] = v[k] << LOG2_CELL_DIM;
        b1[k] = b0[k] + CELL_DIM;
    }
    for (;;) {
        uint32_t const node = g->nodes[index];
        if (!(node & 1u)) {
            break;
        }
        index = node >> 1;
        for (int k = 0; k < 3; ++k) {
            int32_t const center = b0[k] + ((b1[k] - b0[k]) >> 1);
            if (c[k] < center) {
                b1[k] = center;
            } else {
                b0[k] = center;
                index += 1u << k;
            }
        }
    }
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        float const inv = 1.f / (float)m->dims[k];
        lo[k]           = (float)b0[k] * inv;
        hi[k]           = (float)b1[k] * inv;
    }
    float hit_t;
    if (intersect_inside(lo, hi, r, &hit_t)) {
        if (r->max_t > hit_t) {
            r->max_t = hit_t;
        }
    } else {
        r->max_t = r->min_t;
        return 0;
    }
    uint32_t const node = g->nodes[index];
    if (GRID_EMPTY == (node >> 1)) {
        return 0;
    }
    memcpy(cm, g->data + 4 * (size_t)(node >> 1), 4 * sizeof(float));
    return 1;
}

float oracle_similarity_scale(float g, uint32_t depth) {
    float gs;
    if (depth < SR_LOW) {
        gs = g;
    } else if (depth < SR_HIGH) {
        float const x = (1.f / (float)(SR_HIGH - SR_LOW)) * (float)(depth - SR_LOW);
        gs            = (1.f - x) * g + x * 0.f;
    } else {
        gs = 0.f;
    }
    return (1.f - g) / (1.f - gs);
}

/* The collision coefficients per unit density of the nc channels a path
 * carries: the medium's rgb, or a hero-wavelength path's four wavelengths. */
typedef struct {
    int32_t     nc;
    float const* a;
    float const* s;
} channels;

static void coefficients(oracle_medium const* m, channels const* cc, tray const* r, float t,
                         float srs, int32_t nearest, float* a, float* s) {
    float uvw[3];
    tray_point(r, t, uvw);
    float const dn = oracle_medium_density(m, uvw, nearest);
    for (int c = 0; c < cc->nc; ++c) {
        a[c] = dn * cc->a[c];
        s[c] = (dn * cc->s[c]) * srs;
    }
}

static float average(int32_t nc, float const* v) {
    if (3 == nc) {
        return (v[0] + v[1] + v[2]) / 3.f;
    }
    return (((v[0] + v[1]) + v[2]) + v[3]) * 0.25f;
}

static int all_below(int32_t nc, float const* w) {
    for (int c = 0; c < nc; ++c) {
        if (!(w[c] < ABORT_EPSILON)) {
            return 0;
        }
    }
    return 1;
}

/* Tracking::tracking; 1 = scattered at *t_out */
static int tracking(oracle_medium const* m, channels const* cc, tray const* r, float const cm[4],
                    float srs, int32_t nearest, oracle_random_fn rnd, void* rs, float* t_out,
                    float* w) {
    int32_t const nc = cc->nc;
    float const   mt = cm[2] + cm[3];
    if (mt < MIN_MT) {
        return 0;
    }
    float lw[4];
    memcpy(lw, w, sizeof(float) * (size_t)nc);
    float const imt = 1.f / mt;
    for (float t = r->min_t, d = r->max_t;;) {
        float const r0 = rnd(rs);
        t -= logf(1.f - r0) * imt;
        if (t > d) {
            memcpy(w, lw, sizeof(float) * (size_t)nc);
            return 0;
        }
        float mu_a[4], mu_s[4];
        coefficients(m, cc, r, t, srs, nearest, mu_a, mu_s);
        float mu_n[4], sw[4] = {0.f, 0.f, 0.f, 0.f}, nw[4] = {0.f, 0.f, 0.f, 0.f};
        for (int c = 0; c < nc; ++c) {
            mu_n[c] = mt - (mu_a[c] + mu_s[c]);
            sw[c]   = mu_s[c] * lw[c];
            nw[c]   = mu_n[c] * lw[c];
        }
        float const ms = average(nc, sw);
        float const mn = average(nc, nw);
        float const ic = 1.f / (ms + mn);
        float const ps = ms * ic;
        float const pn = mn * ic;
        float const r1 = rnd(rs);
        if ((r1 <= 1.f - pn) & (ps > 0.f)) {
            for (int c = 0; c < nc; ++c) {
                w[c] = lw[c] * (mu_s[c] / (mt * ps));
            }
            *t_out = t;
            return 1;
        }
        for (int c = 0; c < nc; ++c) {
            lw[c] *= mu_n[c] / (mt * pn);
        }
    }
}

/* tracking_transmitted with residual ratio tracking; 0 = below the abort threshold */
static int tracking_transmitted(oracle_medium const* m, channels const* cc, tray const* r,
                                float const cm[4], float srs, int32_t nearest,
                                oracle_random_fn rnd, void* rs, float* w) {
    int32_t const nc = cc->nc;
    float const   mt = cm[2] + cm[3];
    if (mt < MIN_MT) {
        return 1;
    }
    float const minorant = cm[0] + cm[1];
    float       control  = 0.f;
    float       mr       = mt;
    if (minorant > 0.f) {
        float const att = expf(-(r->max_t - r->min_t) * minorant);
        for (int c = 0; c < nc; ++c) {
            w[c] *= att;
        }
        if (all_below(nc, w)) {
            return 0;
        }
        mr = mt - minorant;
        if (mr < MIN_MT) {
            return 1;
        }
        control = minorant;
    }
    float const imt = 1.f / mr;
    for (float t = r->min_t, d = r->max_t;;) {
        float const r0 = rnd(rs);
        t -= logf(1.f - r0) * imt;
        if (t > d) {
            return 1;
        }
        float mu_a[4], mu_s[4];
        coefficients(m, cc, r, t, srs, nearest, mu_a, mu_s);
        for (int c = 0; c < nc; ++c) {
            float const mu_t = (mu_a[c] + mu_s[c]) - control;
            float const mu_n = mr - mu_t;
            w[c] *= imt * mu_n;
        }
        if (all_below(nc, w)) {
            return 0;
        }
    }
}

static void texture_ray(oracle_medium const* m, float const o[3], float const d[3], float t0,
                        float t1, tray* r) {
    for (int k = 0; k < 3; ++k) {
        float const iext = 1.f / (m->box_hi[k] - m->box_lo[k]);
        r->o[k]          = (o[k] - m->box_lo[k]) * iext;
        r->d[k]          = d[k] * iext;
        r->inv[k]        = reciprocal(r->d[k]);
    }
    r->min_t = t0;
    r->max_t = t1;
}

int32_t oracle_medium_overlap(oracle_medium const* m, float const o[3], float const d[3],
                              float min_t, float max_t, float* t0, float* t1) {
    float a = min_t;
    float b = max_t;
    for (int k = 0; k < 3; ++k) {
        /* the box's upper corner as the device keeps it: lo + 1 / (1 / extent) */
        float const iext = 1.f / (m->box_hi[k] - m->box_lo[k]);
        float const lo   = m->box_lo[k];
        float const hi   = m->box_lo[k] + 1.f / iext;
        if (0.f == d[k]) {
            if (o[k] < lo || o[k] > hi) {
                return 0;
            }
            continue;
        }
        float const inv = reciprocal(d[k]);
        float       n   = (lo - o[k]) * inv;
        float       f   = (hi - o[k]) * inv;
        if (n > f) {
            float const x = n;
            n             = f;
            f             = x;
        }
        a = n > a ? n : a;
        b = f < b ? f : b;
    }
    /* a ray from outside steps past the entry as through a boundary surface */
    *t0 = a > min_t ? oracle_offset_f(a) : a;
    *t1 = b;
    return *t0 < *t1;
}

int32_t oracle_medium_track_n(oracle_medium const* m, float const o[3], float const d[3],
                              float t0, float t1, uint32_t depth, int32_t nearest,
                              oracle_random_fn rnd, void* rs, int32_t nc, float const* cc_a,
                              float const* cc_s, float* t_out, float* w) {
    channels const cc = {nc, cc_a, cc_s};
    tray           r;
    texture_ray(m, o, d, t0, t1, &r);
    float const srs   = oracle_similarity_scale(m->g, depth);
    int32_t     event = 0;
    for (int c = 0; c < nc; ++c) {
        w[c] = 1.f;
    }
    for (float const end = t1; r.min_t < end;) {
        float cm[4];
        if (gridtree_intersect(m, &r, cm)) {
            cm[1] *= srs;
            cm[3] *= srs;
            if (tracking(m, &cc, &r, cm, srs, nearest, rnd, rs, t_out, w)) {
                event = 1;
                break;
            }
        }
        r.min_t = oracle_offset_f(r.max_t);
        r.max_t = end;
    }
    return all_below(nc, w) ? 2 : event;
}

int32_t oracle_medium_track(oracle_medium const* m, float const o[3], float const d[3], float t0,
                            float t1, uint32_t depth, int32_t nearest, oracle_random_fn rnd,
                            void* rs, float* t_out, float w[3]) {
    return oracle_medium_track_n(m, o, d, t0, t1, depth, nearest, rnd, rs, 3, m->cc_a, m->cc_s,
                                 t_out, w);
}

int32_t oracle_medium_transmittance_n(oracle_medium const* m, float const o[3], float const d[3],
                                      float t0, float t1, uint32_t depth, int32_t nearest,
                                      oracle_random_fn rnd, void* rs, int32_t nc,
                                      float const* cc_a, float const* cc_s, float* tr) {
    channels const cc = {nc, cc_a, cc_s};
    for (int c = 0; c < nc; ++c) {
        tr[c] = 1.f;
    }
    if (oracle_offset_f(t0) >= t1) {
        return 1;
    }
    tray r;
    texture_ray(m, o, d, t0, t1, &r);
    float const srs = oracle_similarity_scale(m->g, depth);
    for (float const end = t1; r.min_t < end;) {
        float cm[4];
        if (gridtree_intersect(m, &r, cm)) {
            cm[1] *= srs;
            cm[3] *= srs;
            if (!tracking_transmitted(m, &cc, &r, cm, srs, nearest, rnd, rs, tr)) {
                for (int c = 0; c < nc; ++c) {
                    tr[c] = 0.f;
                }
                return 0;
            }
        }
        r.min_t = oracle_offset_f(r.max_t);
        r.max_t = end;
    }
    return 1;
}

int32_t oracle_medium_transmittance(oracle_medium const* m, float const o[3], float const d[3],
                                    float t0, float t1, uint32_t depth, int32_t nearest,
                                    oracle_random_fn rnd, void* rs, float tr[3]) {
    return oracle_medium_transmittance_n(m, o, d, t0, t1, depth, nearest, rnd, rs, 3, m->cc_a,
                                         m->cc_s, tr);
}

float oracle_phase_hg(float cos_theta, float g) {
    float const gg    = g * g;
    float const denom = 1.f + gg + 2.f * g * cos_theta;
    return (1.f / (4.f * PI_F)) * (1.f - gg) / (denom * sqrtf(denom));
}

void oracle_sample_phase(float const wo[3], float r0, float r1, float g, float wi[3], float* pdf) {
    float cos_theta;
    if (fabsf(g) < 0.001f) {
        cos_theta = 1.f - 2.f * r0;
    } else {
        float const gg  = g * g;
        float const sqr = (1.f - gg) / (1.f - g + 2.f * g * r0);
        cos_theta       = (1.f + gg - sqr * sqr) / (2.f * g);
    }
    float const sin_theta = sqrtf(fmaxf(0.f, 1.f - cos_theta * cos_theta));
    float const phi       = r1 * (2.f * PI_F);
    float       t[3], b[3];
    oracle_orthonormal_basis(wo, t, b);
    float const sp = sinf(phi);
    float const cp = cosf(phi);
    for (int k = 0; k < 3; ++k) {
        wi[k] = (sin_theta * cp * t[k] + sin_theta * sp * b[k]) + cos_theta * -wo[k];
    }
    *pdf = oracle_phase_hg(-cos_theta, g);
}

int32_t oracle_medium_batch(oracle_medium const* m, float const* rays, uint32_t n, uint32_t depth,
                            uint64_t seed, int32_t transmittance, int32_t nearest, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        float const* r = rays + 8 * (size_t)i;
        oracle_pcg   g;
        oracle_pcg_start(&g, seed, i);
        float res[5] = {0.f, 0.f, 1.f, 1.f, 1.f};
        float t0, t1;
        if (oracle_medium_overlap(m, r, r + 4, r[3], r[7], &t0, &t1)) {
            if (transmittance) {
                res[0] = oracle_medium_transmittance(m, r, r + 4, t0, t1, depth, nearest,
                                                     oracle_pcg_float, &g, res + 2)
                             ? 1.f
                             : 0.f;
            } else {
                float t_out = 0.f;
                res[0] = (float)oracle_medium_track(m, r, r + 4, t0, t1, depth, nearest,
                                                    oracle_pcg_float, &g, &t_out, res + 2);
                res[1] = t_out;
            }
        } else if (transmittance) {
            res[0] = 1.f;
        }
        memcpy(out + 5 * (size_t)i, res, sizeof(res));
    }
    return 0;
}
