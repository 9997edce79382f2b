/*
 * ORACLE (test infrastructure only; never linked into the product).
 * CPU restatement of sprout's two-level traversal of instanced meshes:
 *
 *   BVH_wrapper::intersect / intersect_p   core/scene/prop/prop_bvh_wrapper.cpp:33-106, 183-247
 *   Prop::intersect / intersect_p          core/scene/prop/prop.cpp:68-126
 *   Mesh::intersect / intersect_p          core/scene/shape/triangle/triangle_mesh.cpp:366-446
 *   Tree::intersect on the shared stack    core/scene/shape/triangle/bvh/triangle_bvh_tree.cpp:41-90
 *   AABB::transform                        base/math/aabb.inl:263-280
 *   transform_point / transform_vector     base/math/matrix4x4.inl:319-342, as the
 *                                          reference build compiles them:
 *                                          fma(z, r2, fma(x, r0, y * r1)) (+ r3)
 *   prop tree: scene::bvh::Builder         core/scene/bvh/scene_bvh_builder.cpp:14-107,
 *                                          the same split as the mesh builder over the
 *                                          instances' world boxes
 */

#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SENTINEL 0xFFFFFFFFu
#define STACK_SIZE 128

void oracle_transform_point(float const m[16], float const v[3], float out[3]) {
    for (int c = 0; c < 3; ++c) {
        out[c] = fmaf(v[2], m[8 + c], fmaf(v[0], m[0 + c], v[1] * m[4 + c])) + m[12 + c];
    }
}

void oracle_transform_vector(float const m[16], float const v[3], float out[3]) {
    for (int c = 0; c < 3; ++c) {
        out[c] = fmaf(v[2], m[8 + c], fmaf(v[0], m[0 + c], v[1] * m[4 + c]));
    }
}

static float fmin2(float a, float b) {
    return b < a ? b : a; /* std::min */
}
static float fmax2(float a, float b) {
    return a < b ? b : a; /* std::max */
}

void oracle_transform_box(float const lo[3], float const hi[3], float const m[16], float out_lo[3],
                          float out_hi[3]) {
    for (int c = 0; c < 3; ++c) {
        float const xa = lo[0] * m[0 + c], xb = hi[0] * m[0 + c];
        float const ya = lo[1] * m[4 + c], yb = hi[1] * m[4 + c];
        float const za = lo[2] * m[8 + c], zb = hi[2] * m[8 + c];
        float const w  = m[12 + c];
        out_lo[c]      = (fmin2(xa, xb) + fmin2(ya, yb)) + (fmin2(za, zb) + w);
        out_hi[c]      = (fmax2(xa, xb) + fmax2(ya, yb)) + (fmax2(za, zb) + w);
    }
}

int32_t oracle_instances_build(oracle_mesh const* meshes, uint32_t num_meshes,
                               uint32_t const* mesh_of, float const* world_to_object,
                               float const* object_to_world, uint32_t num_instances,
                               oracle_instances* out) {
    memset(out, 0, sizeof(*out));
    if (!meshes || !mesh_of || !world_to_object || !object_to_world || 0 == num_instances) {
        return ORACLE_ERR_ARGS;
    }
    out->boxes = (float*)malloc(sizeof(float) * 6 * num_instances);
    float*    pos = (float*)malloc(sizeof(float) * 6 * num_instances);
    uint32_t* idx = (uint32_t*)malloc(sizeof(uint32_t) * 3 * num_instances);
    if (!out->boxes || !pos || !idx) {
        free(pos);
        free(idx);
        free(out->boxes);
        out->boxes = NULL;
        return ORACLE_ERR_ALLOC;
    }
    for (uint32_t i = 0; i < num_instances; ++i) {
        if (mesh_of[i] >= num_meshes) {
            free(pos);
            free(idx);
            free(out->boxes);
            out->boxes = NULL;
            return ORACLE_ERR_ARGS;
        }
        onode const* root = &meshes[mesh_of[i]].bvh->nodes[0];
        float*       b    = out->boxes + 6 * (size_t)i;
        oracle_transform_box(root->lo, root->hi, object_to_world + 16 * (size_t)i, b, b + 3);
        memcpy(pos + 6 * (size_t)i, b, 6 * sizeof(float));
        idx[3 * (size_t)i + 0] = 2 * i;
        idx[3 * (size_t)i + 1] = 2 * i + 1;
        idx[3 * (size_t)i + 2] = 2 * i;
    }
    /* the box (lo, hi, lo) of each degenerate triangle is the instance's box */
    int32_t const st = oracle_sbvh_build(pos, 2 * num_instances, idx, num_instances, 16, 64, 4, &out->top);
    free(pos);
    free(idx);
    if (st) {
        free(out->boxes);
        out->boxes = NULL;
        return st;
    }
    out->w2o           = world_to_object;
    out->mesh_of       = mesh_of;
    out->meshes        = meshes;
    out->num_instances = num_instances;
    return 0;
}

void oracle_instances_free(oracle_instances* t) {
    oracle_bvh_free(&t->top);
    free(t->boxes);
    t->boxes = NULL;
}

static float smin(float a, float b) {
    return a < b ? a : b;
}
static float smax(float a, float b) {
    return a > b ? a : b;
}

static int slab(float const lo[3], float const hi[3], float const o[3], float const inv[3],
                float min_t, float max_t) {
    float tmax_c[3], tmin_c[3];
    for (int i = 0; i < 3; ++i) {
        float const l1 = (lo[i] - o[i]) * inv[i];
        float const l2 = (hi[i] - o[i]) * inv[i];
        tmax_c[i]      = smax(smin(l1, __builtin_inff()), smin(l2, __builtin_inff()));
        tmin_c[i]      = smin(smax(l1, -__builtin_inff()), smax(l2, -__builtin_inff()));
    }
    float const tmax = smin(smin(tmax_c[0], tmax_c[1]), tmax_c[2]);
    float const tmin = smax(smax(tmin_c[0], tmin_c[1]), tmin_c[2]);
    return (tmax >= min_t) & (max_t >= tmin) & (tmax >= tmin);
}

int32_t oracle_instances_trace(oracle_instances const* t, float const* rays, uint64_t num_rays,
                               int32_t any, uint32_t* hits) {
    uint32_t stack[STACK_SIZE];
    for (uint64_t r = 0; r < num_rays; ++r) {
        float const* ray = rays + 8 * r;
        float const  o[3] = {ray[0], ray[1], ray[2]};
        float const  d[3] = {ray[4], ray[5], ray[6]};
        float const  min_t = ray[3];
        float        max_t = ray[7];
        float        inv[3];
        oracle_reciprocal3(d, ORACLE_RCP_PINNED, inv);
        uint32_t const sign[3] = {inv[0] < 0.f, inv[1] < 0.f, inv[2] < 0.f};

        uint32_t instance = SENTINEL, index = SENTINEL;
        float    u = 0.f, v = 0.f;
        int      hit_any = 0;

        uint32_t sp = 0;
        stack[sp++] = SENTINEL;
        uint32_t n  = t->top.num_nodes ? 0u : SENTINEL;
        while (SENTINEL != n) {
            onode const* node = &t->top.nodes[n];
            if (slab(node->lo, node->hi, o, inv, min_t, max_t)) {
                if (0 == node->num) {
                    if (sp >= STACK_SIZE) {
                        return ORACLE_ERR_STACK;
                    }
                    if (0 == sign[node->axis]) {
                        stack[sp++] = node->data + 1;
                        n           = node->data;
                    } else {
                        stack[sp++] = node->data;
                        n           = node->data + 1;
                    }
                    continue;
                }
                for (uint32_t i = node->data, end = node->data + node->num; i < end; ++i) {
                    uint32_t const p   = t->top.slots[i];
                    float const*   box = t->boxes + 6 * (size_t)p;
                    if (!slab(box, box + 3, o, inv, min_t, max_t)) {
                        continue;
                    }
                    float const* m = t->w2o + 16 * (size_t)p;
                    float        oo[3], od[3], oinv[3];
                    oracle_transform_point(m, o, oo);
                    oracle_transform_vector(m, d, od);
                    oracle_reciprocal3(od, ORACLE_RCP_PINNED, oinv);
                    uint32_t const osign[3] = {oinv[0] < 0.f, oinv[1] < 0.f, oinv[2] < 0.f};

                    oracle_mesh const* mesh = &t->meshes[t->mesh_of[p]];
                    /* the mesh's walk on the same stack, above a sentinel */
                    if (sp >= STACK_SIZE) {
                        return ORACLE_ERR_STACK;
                    }
                    stack[sp++] = SENTINEL;
                    uint32_t mn = 0;
                    float    omax = max_t;
                    uint32_t midx = SENTINEL;
                    float    mu = 0.f, mv = 0.f;
                    int      mhit = 0;
                    while (SENTINEL != mn) {
                        onode const* mnode = &mesh->bvh->nodes[mn];
                        if (slab(mnode->lo, mnode->hi, oo, oinv, min_t, omax)) {
                            if (0 == mnode->num) {
                                if (sp >= STACK_SIZE) {
                                    return ORACLE_ERR_STACK;
                                }
                                if (0 == osign[mnode->axis]) {
                                    stack[sp++] = mnode->data + 1;
                                    mn          = mnode->data;
                                } else {
                                    stack[sp++] = mnode->data;
                                    mn          = mnode->data + 1;
                                }
                                continue;
                            }
                            for (uint32_t k = mnode->data, ke = mnode->data + mnode->num; k < ke; ++k) {
                                uint32_t const tri = mesh->bvh->slots[k];
                                float const* pa = mesh->positions + 3 * (size_t)mesh->indices[3 * tri + 0];
                                float const* pb = mesh->positions + 3 * (size_t)mesh->indices[3 * tri + 1];
                                float const* pc = mesh->positions + 3 * (size_t)mesh->indices[3 * tri + 2];
                                if (any) {
                                    float tmax = omax, tu, tv;
                                    if (oracle_triangle(pa, pb, pc, oo, od, min_t, &tmax, &tu, &tv, ORACLE_RCP_PINNED)) {
                                        mhit = 1;
                                        break;
                                    }
                                } else if (oracle_triangle(pa, pb, pc, oo, od, min_t, &omax, &mu, &mv, ORACLE_RCP_PINNED)) {
                                    midx = tri;
                                    mhit = 1;
                                }
                            }
                            if (any && mhit) {
                                break;
                            }
                        }
                        mn = stack[--sp];
                    }
                    if (any && mhit) {
                        hit_any = 1;
                        break;
                    }
                    if (mhit) {
                        max_t    = omax;
                        instance = p;
                        index    = midx;
                        u        = mu;
                        v        = mv;
                    }
                }
                if (hit_any) {
                    break;
                }
            }
            n = stack[--sp];
        }

        if (any) {
            hits[r] = (uint32_t)hit_any;
        } else {
            uint32_t* h = hits + 5 * r;
            memcpy(h + 0, &max_t, 4);
            if (SENTINEL != index) {
                memcpy(h + 1, &u, 4);
                memcpy(h + 2, &v, 4);
            } else {
                h[1] = h[2] = 0;
            }
            h[3] = index;
            h[4] = instance;
        }
    }
    return 0;
}
