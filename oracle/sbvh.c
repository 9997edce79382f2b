/*
 * ORACLE (test infrastructure only; never linked into the product).
 *
 * CPU restatement of sprout's triangle SBVH builder:
 *   scene::bvh::Kernel::split            core/scene/bvh/scene_bvh_builder_base.cpp:26-88
 *   scene::bvh::Kernel::splitting_plane  core/scene/bvh/scene_bvh_builder_base.cpp:90-179
 *   scene::bvh::Kernel::assign           core/scene/bvh/scene_bvh_builder_base.cpp:181-189
 *   scene::bvh::Builder_base::split      core/scene/bvh/scene_bvh_builder_base.cpp:271-290
 *   scene::bvh::Split_candidate          core/scene/bvh/scene_bvh_split_candidate.inl:49-151
 *   triangle Builder_SAH::build/serialize core/scene/shape/triangle/bvh/triangle_bvh_builder_sah.cpp:21-67,129-164
 *   scene::bvh::Node (32-byte layout)    core/scene/bvh/scene_bvh_node.hpp:44-59
 *
 * The reference runs sub-trees as thread-pool tasks; the tree it produces does
 * not depend on that (a task re-runs the same serial split on its sub-tree and
 * serialize() re-lays the nodes out depth-first), so this restatement is serial.
 *
 * Floating point follows the reference as compiled by g++ -O3 -march=native,
 * which contracts some products into FMAs.  Those contractions are spelled out
 * here with fmaf() and this file is compiled with -ffp-contract=off:
 *   surface area  2 * fma(d1, d2, fma(d0, d1, d0 * d2))      base/math/aabb.inl:39-42
 *   slice plane   fma(i, step_a, min[a])                       scene_bvh_builder_base.cpp:133
 *   SAH cost      fma(k, 0.125, fma(n0, SA0, n1 * SA1) / SA + 2)  scene_bvh_split_candidate.inl:109-118
 * min/max follow SSE minps/maxps: min(a, b) = a < b ? a : b.
 */

#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    float    lo[3];
    uint32_t prim;
    float    hi[3];
    uint32_t pad;
} oref;

typedef struct {
    float lo[3];
    float hi[3];
} obox;

typedef struct {
    obox     b0;
    obox     b1;
    uint32_t n0;
    uint32_t n1;
    float    d;
    float    cost;
    uint8_t  axis;
    uint8_t  spatial;
} ocand;

typedef struct {
    oref*    v;
    uint32_t n;
} oreflist;

typedef struct {
    uint32_t num_slices;
    uint32_t sweep_threshold;
    uint32_t max_primitives;
    uint32_t spatial_split_threshold;

    onode*   nodes; /* build order */
    uint32_t num_nodes;
    uint32_t cap_nodes;

    uint32_t* ref_ids;
    uint32_t  num_ref_ids;
    uint32_t  cap_ref_ids;

    ocand*   cands;
    uint32_t cap_cands;

    int32_t status;
} obuilder;

static float smin(float a, float b) {
    return a < b ? a : b;
}

static float smax(float a, float b) {
    return a > b ? a : b;
}

/* std::min / std::max */
static float stdmin(float a, float b) {
    return b < a ? b : a;
}

static float stdmax(float a, float b) {
    return a < b ? b : a;
}

static obox empty_box(void) {
    obox b;
    for (int i = 0; i < 3; ++i) {
        b.lo[i] = 3.40282347e+38f;
        b.hi[i] = -3.40282347e+38f;
    }
    return b;
}

static void merge_ref(obox* b, oref const* r) {
    for (int i = 0; i < 3; ++i) {
        b->lo[i] = smin(b->lo[i], r->lo[i]);
        b->hi[i] = smax(b->hi[i], r->hi[i]);
    }
}

/* base/math/aabb.inl:39-42 as contracted by the reference build */
float oracle_surface_area(float const lo[3], float const hi[3]) {
    float const d0 = hi[0] - lo[0];
    float const d1 = hi[1] - lo[1];
    float const d2 = hi[2] - lo[2];
    float const s  = fmaf(d1, d2, fmaf(d0, d1, d0 * d2));
    return s + s;
}

static int covers(obox const* a, obox const* b) {
    return a->lo[0] <= b->lo[0] && a->lo[1] <= b->lo[1] && a->lo[2] <= b->lo[2] &&
           a->hi[0] >= b->hi[0] && a->hi[1] >= b->hi[1] && a->hi[2] >= b->hi[2];
}

static void push_node(obuilder* b) {
    if (b->num_nodes == b->cap_nodes) {
        b->cap_nodes = b->cap_nodes ? 2 * b->cap_nodes : 1024;
        b->nodes     = (onode*)realloc(b->nodes, sizeof(onode) * b->cap_nodes);
    }
    memset(&b->nodes[b->num_nodes], 0, sizeof(onode));
    ++b->num_nodes;
}

static void set_aabb(onode* n, obox const* box) {
    for (int i = 0; i < 3; ++i) {
        n->lo[i] = box->lo[i];
        n->hi[i] = box->hi[i];
    }
}

static void assign(obuilder* b, uint32_t node_id, oreflist const* refs) {
    onode* n = &b->nodes[node_id];

    n->data = b->num_ref_ids;
    n->num  = (uint8_t)refs->n;

    if (b->num_ref_ids + refs->n > b->cap_ref_ids) {
        while (b->num_ref_ids + refs->n > b->cap_ref_ids) {
            b->cap_ref_ids = b->cap_ref_ids ? 2 * b->cap_ref_ids : 4096;
        }
        b->ref_ids = (uint32_t*)realloc(b->ref_ids, sizeof(uint32_t) * b->cap_ref_ids);
    }

    for (uint32_t i = 0; i < refs->n; ++i) {
        b->ref_ids[b->num_ref_ids++] = refs->v[i].prim;
    }
}

static void cand_init(ocand* c, uint8_t axis, float const p[3], uint8_t spatial) {
    c->b0      = empty_box();
    c->b1      = empty_box();
    c->d       = p[axis];
    c->axis    = axis;
    c->spatial = spatial;
}

/* Split_candidate::evaluate, scene_bvh_split_candidate.inl:49-123 */
static void cand_evaluate(ocand* c, oreflist const* refs, float aabb_surface_area) {
    uint32_t n0 = 0;
    uint32_t n1 = 0;

    uint8_t const a = c->axis;
    float const   d = c->d;

    if (c->spatial) {
        int used_spatial = 0;

        for (uint32_t i = 0; i < refs->n; ++i) {
            oref const* r = &refs->v[i];

            if (r->hi[a] < d) {
                ++n0;
                merge_ref(&c->b0, r);
            } else if (!(r->lo[a] < d)) {
                ++n1;
                merge_ref(&c->b1, r);
            } else {
                ++n0;
                ++n1;
                merge_ref(&c->b0, r);
                merge_ref(&c->b1, r);
                used_spatial = 1;
            }
        }

        if (used_spatial) {
            c->b0.hi[a] = stdmin(d, c->b0.hi[a]);
            c->b1.lo[a] = stdmax(d, c->b1.lo[a]);
        } else {
            c->spatial = 0;
        }
    } else {
        for (uint32_t i = 0; i < refs->n; ++i) {
            oref const* r = &refs->v[i];

            if (r->hi[a] < d) {
                ++n0;
                merge_ref(&c->b0, r);
            } else {
                ++n1;
                merge_ref(&c->b1, r);
            }
        }
    }

    if (0 == n0 || 0 == n1) {
        c->cost = 2.f + (float)refs->n;
    } else {
        float const sa0 = oracle_surface_area(c->b0.lo, c->b0.hi);
        float const sa1 = oracle_surface_area(c->b1.lo, c->b1.hi);

        float const weight_1 = (float)n1 * sa1;
        float const weights  = fmaf((float)n0, sa0, weight_1);

        float const k = (float)(uint32_t)(n0 + n1 - refs->n);

        c->cost = fmaf(k, 0.125f, weights / aabb_surface_area + 2.f);
    }

    c->n0 = n0;
    c->n1 = n1;
}

static uint32_t index_max_component(float const v[3]) {
    if (v[0] > v[1]) {
        return v[0] > v[2] ? 0 : 2;
    }
    return v[1] > v[2] ? 1 : 2;
}

/* Kernel::splitting_plane, scene_bvh_builder_base.cpp:90-179.  Returns the
 * index of the chosen candidate in b->cands. */
static uint32_t splitting_plane(obuilder* b, oreflist const* refs, obox const* box,
                                uint32_t depth, int* exhausted) {
    uint32_t nc = 0;

    float position[3];
    float extent[3];
    for (int i = 0; i < 3; ++i) {
        position[i] = (box->lo[i] + box->hi[i]) * 0.5f;
        extent[i]   = box->hi[i] - box->lo[i];
    }

    uint32_t const need = 3 + 3 * (refs->n <= b->sweep_threshold ? refs->n : 2 * (b->num_slices + 2));
    if (need > b->cap_cands) {
        b->cap_cands = need;
        b->cands     = (ocand*)realloc(b->cands, sizeof(ocand) * need);
    }

    cand_init(&b->cands[nc++], 0, position, 1);
    cand_init(&b->cands[nc++], 1, position, 1);
    cand_init(&b->cands[nc++], 2, position, 1);

    if (refs->n <= b->sweep_threshold) {
        for (uint32_t i = 0; i < refs->n; ++i) {
            float const* mx = refs->v[i].hi;
            cand_init(&b->cands[nc++], 0, mx, 0);
            cand_init(&b->cands[nc++], 1, mx, 0);
            cand_init(&b->cands[nc++], 2, mx, 0);
        }
    } else {
        uint32_t const la   = index_max_component(extent);
        float const    step = extent[la] / (float)b->num_slices;

        for (uint8_t a = 0; a < 3; ++a) {
            float const    extent_a  = extent[a];
            uint32_t const num_steps = (uint32_t)ceilf(extent_a / step);
            float const    step_a    = extent_a / (float)num_steps;

            for (uint32_t i = 1; i < num_steps; ++i) {
                float slice[3] = {position[0], position[1], position[2]};
                slice[a]       = fmaf((float)i, step_a, box->lo[a]);

                if (nc + 2 > b->cap_cands) {
                    b->cap_cands = 2 * b->cap_cands;
                    b->cands     = (ocand*)realloc(b->cands, sizeof(ocand) * b->cap_cands);
                }

                cand_init(&b->cands[nc++], a, slice, 0);
                if (depth < b->spatial_split_threshold) {
                    cand_init(&b->cands[nc++], a, slice, 1);
                }
            }
        }
    }

    float const sa = oracle_surface_area(box->lo, box->hi);

    for (uint32_t i = 0; i < nc; ++i) {
        cand_evaluate(&b->cands[i], refs, sa);
    }

    uint32_t best     = 0;
    float    min_cost = b->cands[0].cost;
    for (uint32_t i = 1; i < nc; ++i) {
        if (b->cands[i].cost < min_cost) {
            best     = i;
            min_cost = b->cands[i].cost;
        }
    }

    ocand const* sp = &b->cands[best];

    *exhausted = (covers(&sp->b0, box) && refs->n == sp->n0) ||
                 (covers(&sp->b1, box) && refs->n == sp->n1);

    return best;
}

/* Split_candidate::distribute, scene_bvh_split_candidate.inl:125-151 */
static void distribute(ocand const* c, oreflist const* refs, oreflist* r0, oreflist* r1) {
    r0->v = (oref*)malloc(sizeof(oref) * (c->n0 ? c->n0 : 1));
    r1->v = (oref*)malloc(sizeof(oref) * (c->n1 ? c->n1 : 1));
    r0->n = 0;
    r1->n = 0;

    uint8_t const a = c->axis;
    float const   d = c->d;

    for (uint32_t i = 0; i < refs->n; ++i) {
        oref const* r = &refs->v[i];

        if (r->hi[a] < d) {
            r0->v[r0->n++] = *r;
        } else if (!c->spatial || !(r->lo[a] < d)) {
            r1->v[r1->n++] = *r;
        } else {
            oref x = *r;
            oref y = *r;
            x.hi[a] = stdmin(d, x.hi[a]);
            y.lo[a] = stdmax(d, y.lo[a]);
            r0->v[r0->n++] = x;
            r1->v[r1->n++] = y;
        }
    }
}

/* Kernel::split, scene_bvh_builder_base.cpp:26-88 */
static void split(obuilder* b, uint32_t node_id, oreflist* refs, obox const* box, uint32_t depth) {
    set_aabb(&b->nodes[node_id], box);

    uint32_t const n = refs->n;

    if (n <= b->max_primitives) {
        assign(b, node_id, refs);
        return;
    }

    int            exhausted;
    uint32_t const best = splitting_plane(b, refs, box, depth, &exhausted);
    ocand const    sp   = b->cands[best];

    if (n <= 0xFF && ((float)n <= sp.cost || exhausted)) {
        assign(b, node_id, refs);
        return;
    }

    if (exhausted) {
        /* The reference logs "Cannot split node further" and leaves the node
         * unassigned (scene_bvh_builder_base.cpp:50-57). */
        b->status = ORACLE_ERR_UNSPLITTABLE;
        return;
    }

    oreflist r0;
    oreflist r1;
    distribute(&sp, refs, &r0, &r1);

    if (n <= 0xFF && (0 == r0.n || 0 == r1.n)) {
        free(r0.v);
        free(r1.v);
        assign(b, node_id, refs);
        return;
    }

    ++depth;

    free(refs->v);
    refs->v = NULL;
    refs->n = 0;

    uint32_t const child0 = b->num_nodes;

    b->nodes[node_id].data = child0;
    b->nodes[node_id].axis = sp.axis;
    b->nodes[node_id].num  = 0;

    push_node(b);
    push_node(b);

    split(b, child0, &r0, &sp.b0, depth);
    free(r0.v);

    split(b, child0 + 1, &r1, &sp.b1, depth);
    free(r1.v);
}

typedef struct {
    onode const*    src;
    uint32_t const* ref_ids;
    onode*          dst;
    uint32_t*       slots;
    uint32_t        current_node;
    uint32_t        current_slot;
} oserial;

/* Builder_SAH::serialize, triangle_bvh_builder_sah.cpp:129-164 */
static void serialize(oserial* s, uint32_t source, uint32_t dest) {
    onode const* node = &s->src[source];
    onode*       n    = &s->dst[dest];

    memset(n, 0, sizeof(onode));
    memcpy(n->lo, node->lo, sizeof(n->lo));
    memcpy(n->hi, node->hi, sizeof(n->hi));

    if (0 == node->num) {
        uint32_t const child0 = s->current_node;
        n->data               = child0;
        n->axis               = node->axis;
        n->num                = 0;
        s->current_node += 2;
        serialize(s, node->data, child0);
        serialize(s, node->data + 1, child0 + 1);
    } else {
        n->data = s->current_slot;
        n->num  = node->num;
        for (uint32_t p = node->data, end = node->data + node->num; p < end; ++p) {
            s->slots[s->current_slot++] = s->ref_ids[p];
        }
    }
}

int32_t oracle_sbvh_build(float const* positions, uint32_t num_vertices, uint32_t const* indices,
                          uint32_t num_triangles, uint32_t num_slices, uint32_t sweep_threshold,
                          uint32_t max_primitives, oracle_bvh* out) {
    memset(out, 0, sizeof(*out));

    if (0 == num_triangles || 0 == num_slices || 0 == max_primitives) {
        return ORACLE_ERR_ARGS;
    }

    for (uint32_t i = 0; i < 3 * num_triangles; ++i) {
        if (indices[i] >= num_vertices) {
            return ORACLE_ERR_ARGS;
        }
    }

    obuilder b;
    memset(&b, 0, sizeof(b));
    b.num_slices      = num_slices;
    b.sweep_threshold = sweep_threshold;
    b.max_primitives  = max_primitives;

    oreflist refs;
    refs.n = num_triangles;
    refs.v = (oref*)malloc(sizeof(oref) * num_triangles);

    obox root = empty_box();

    /* triangle_bvh_builder_sah.cpp:30-51, triangle_bvh_helper.cpp:15-21 */
    for (uint32_t t = 0; t < num_triangles; ++t) {
        float const* pa = positions + 3 * (size_t)indices[3 * t + 0];
        float const* pb = positions + 3 * (size_t)indices[3 * t + 1];
        float const* pc = positions + 3 * (size_t)indices[3 * t + 2];

        oref* r = &refs.v[t];
        for (int i = 0; i < 3; ++i) {
            r->lo[i] = smin(pa[i], smin(pb[i], pc[i]));
            r->hi[i] = smax(pa[i], smax(pb[i], pc[i]));
        }
        r->prim = t;
        r->pad  = 0;

        merge_ref(&root, r);
    }

    /* Builder_base::split, scene_bvh_builder_base.cpp:271-290 */
    float const log2_n        = log2f((float)num_triangles);
    b.spatial_split_threshold = (uint32_t)lrintf(log2_n / 2.f);

    push_node(&b);

    split(&b, 0, &refs, &root, 0);
    free(refs.v);

    if (b.status) {
        free(b.nodes);
        free(b.ref_ids);
        free(b.cands);
        return b.status;
    }

    out->num_nodes = b.num_nodes;
    out->num_slots = b.num_ref_ids;
    out->nodes     = (onode*)malloc(sizeof(onode) * b.num_nodes);
    out->slots     = (uint32_t*)malloc(sizeof(uint32_t) * (b.num_ref_ids ? b.num_ref_ids : 1));
    out->spatial_split_threshold = b.spatial_split_threshold;

    oserial s;
    s.src          = b.nodes;
    s.ref_ids      = b.ref_ids;
    s.dst          = out->nodes;
    s.slots        = out->slots;
    s.current_node = 1;
    s.current_slot = 0;
    serialize(&s, 0, 0);

    free(b.nodes);
    free(b.ref_ids);
    free(b.cands);

    return 0;
}

void oracle_bvh_free(oracle_bvh* bvh) {
    free(bvh->nodes);
    free(bvh->slots);
    memset(bvh, 0, sizeof(*bvh));
}
