/*
 * ORACLE (test infrastructure only; never linked into the product).
 *
 * CPU restatement of sprout's triangle BVH traversal:
 *   Tree::intersect / intersect_p      core/scene/shape/triangle/bvh/triangle_bvh_tree.cpp:41-90,145-185
 *   Node::intersect_p (slab test)      core/scene/bvh/scene_bvh_node.inl:133-163
 *   triangle::intersect / intersect_p  core/scene/shape/triangle/triangle_primitive_mt.inl:61-93,191-218
 *   reciprocal3 / reciprocal           base/math/simd.inl:206-212, base/math/scalar.inl:67-71
 *   cross3 / dot3_scalar               base/math/simd.inl:165-188
 *   sign                               base/math/simd.inl:230-236
 *   Node_stack (127 entries)           core/scene/shape/node_stack.hpp:8-28
 *   offset_ray                         core/scene/ray_offset.inl:18-29
 *
 * Floating point follows the reference as compiled by g++ -O3 -march=native
 * (this file is compiled with -ffp-contract=off; fused steps use fmaf):
 *   reciprocal3(d) = fma(-(r*r), d, r + r)        (the compiler fuses the Newton step)
 *   reciprocal(x)  = (r + r) - (r*r)*x            (scalar intrinsics, not fused)
 *   cross3(a, b).x = fma(a.y, b.z, -(a.z*b.y))    (explicit _mm_fmsub_ps under AVX2), etc.
 *   dot3(a, b)     = (a.x*b.x + a.y*b.y) + a.z*b.z
 * where r is this host's rcpps estimate (NATIVE) or the correctly rounded 1/x
 * (PINNED, the definition the GPU kernels implement).
 */

#include "oracle.h"

#include <immintrin.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define STACK_SIZE 127
#define SENTINEL 0xFFFFFFFFu

static float rcp_estimate(float x, int32_t mode) {
    if (ORACLE_RCP_NATIVE == mode) {
        return _mm_cvtss_f32(_mm_rcp_ss(_mm_set_ss(x)));
    }
    return 1.f / x;
}

float oracle_reciprocal(float x, int32_t mode) {
    float const r  = rcp_estimate(x, mode);
    float const rr = r * r;
    float const m  = rr * x;
    return (r + r) - m;
}

void oracle_reciprocal3(float const d[3], int32_t mode, float out[3]) {
    for (int i = 0; i < 3; ++i) {
        float const r  = rcp_estimate(d[i], mode);
        float const rr = r * r;
        out[i]         = fmaf(-rr, d[i], r + r);
    }
}

static float smin(float a, float b) {
    return a < b ? a : b;
}

static float smax(float a, float b) {
    return a > b ? a : b;
}

static void cross3(float const a[3], float const b[3], float c[3]) {
    c[0] = fmaf(a[1], b[2], -(a[2] * b[1]));
    c[1] = fmaf(a[2], b[0], -(a[0] * b[2]));
    c[2] = fmaf(a[0], b[1], -(a[1] * b[0]));
}

static float dot3(float const a[3], float const b[3]) {
    float const x = a[0] * b[0];
    float const y = a[1] * b[1];
    float const z = a[2] * b[2];
    return (x + y) + z;
}

/* Node::intersect_p */
static int slab(onode const* n, float const o[3], float const inv[3], float min_t, float max_t) {
    float tmax_c[3];
    float tmin_c[3];
    for (int i = 0; i < 3; ++i) {
        float const l1  = (n->lo[i] - o[i]) * inv[i];
        float const l2  = (n->hi[i] - o[i]) * inv[i];
        float const l1a = smin(l1, __builtin_inff());
        float const l2a = smin(l2, __builtin_inff());
        float const l1b = smax(l1, -__builtin_inff());
        float const l2b = smax(l2, -__builtin_inff());
        tmax_c[i]       = smax(l1a, l2a);
        tmin_c[i]       = smin(l1b, l2b);
    }
    float const tmax = smin(smin(tmax_c[0], tmax_c[1]), tmax_c[2]);
    float const tmin = smax(smax(tmin_c[0], tmin_c[1]), tmin_c[2]);
    return (tmax >= min_t) & (max_t >= tmin) & (tmax >= tmin);
}

/* triangle::intersect: returns 1 and updates *max_t, *u, *v on a hit */
static int triangle(float const* pa, float const* pb, float const* pc, float const o[3],
                    float const d[3], float min_t, float* max_t, float* u_out, float* v_out,
                    int32_t mode) {
    float e1[3], e2[3], tv[3];
    for (int i = 0; i < 3; ++i) {
        e1[i] = pb[i] - pa[i];
        e2[i] = pc[i] - pa[i];
        tv[i] = o[i] - pa[i];
    }

    float p[3], q[3];
    cross3(d, e2, p);
    cross3(tv, e1, q);

    float const e1_d_pv = dot3(e1, p);
    float const tv_d_pv = dot3(tv, p);
    float const di_d_qv = dot3(d, q);
    float const e2_d_qv = dot3(e2, q);

    float const inv_det = oracle_reciprocal(e1_d_pv, mode);

    float const u  = tv_d_pv * inv_det;
    float const v  = di_d_qv * inv_det;
    float const t  = e2_d_qv * inv_det;
    float const uv = u + v;

    if ((u >= 0.f) & (1.f >= u) & (v >= 0.f) & (1.f >= uv) & (t >= min_t) & (*max_t >= t)) {
        *max_t = t;
        *u_out = u;
        *v_out = v;
        return 1;
    }
    return 0;
}

int32_t oracle_triangle(float const* pa, float const* pb, float const* pc, float const o[3],
                        float const d[3], float min_t, float* max_t, float* u, float* v,
                        int32_t rcp_mode) {
    return triangle(pa, pb, pc, o, d, min_t, max_t, u, v, rcp_mode);
}

typedef struct {
    onode const*    nodes;
    uint32_t const* slots;
    float const*    positions;
    uint32_t const* indices;
    float const*    rays;
    uint64_t        begin;
    uint64_t        end;
    int32_t         any;
    int32_t         mode;
    uint32_t*       hits;
    uint32_t*       counts;
    int32_t         status;
} otask;

static void trace_range(otask* w) {
    uint32_t stack[STACK_SIZE];

    for (uint64_t r = w->begin; r < w->end; ++r) {
        float const* ray = w->rays + 8 * r;

        float const o[3] = {ray[0], ray[1], ray[2]};
        float const d[3] = {ray[4], ray[5], ray[6]};
        float const min_t = ray[3];
        float       max_t = ray[7];

        float inv[3];
        oracle_reciprocal3(d, w->mode, inv);

        uint32_t const sign[3] = {inv[0] < 0.f, inv[1] < 0.f, inv[2] < 0.f};

        uint32_t sp = 0;
        stack[sp++] = SENTINEL;

        uint32_t n     = 0;
        uint32_t index = SENTINEL;
        float    u     = 0.f;
        float    v     = 0.f;
        int      hit_p = 0;

        uint32_t num_nodes = 0;
        uint32_t num_tris  = 0;

        while (SENTINEL != n) {
            onode const* node = &w->nodes[n];
            ++num_nodes;

            if (slab(node, o, inv, min_t, max_t)) {
                if (0 == node->num) {
                    uint32_t const a = node->data;
                    uint32_t const b = a + 1;

                    if (sp == STACK_SIZE) {
                        w->status = ORACLE_ERR_STACK;
                        break;
                    }

                    if (0 == sign[node->axis]) {
                        stack[sp++] = b;
                        n           = a;
                    } else {
                        stack[sp++] = a;
                        n           = b;
                    }
                    continue;
                }

                for (uint32_t i = node->data, len = node->data + node->num; i < len; ++i) {
                    uint32_t const tri = w->slots[i];

                    float const* pa = w->positions + 3 * (size_t)w->indices[3 * tri + 0];
                    float const* pb = w->positions + 3 * (size_t)w->indices[3 * tri + 1];
                    float const* pc = w->positions + 3 * (size_t)w->indices[3 * tri + 2];

                    ++num_tris;

                    float tu, tv;
                    if (w->any) {
                        float tmax = max_t;
                        if (triangle(pa, pb, pc, o, d, min_t, &tmax, &tu, &tv, w->mode)) {
                            hit_p = 1;
                            break;
                        }
                    } else if (triangle(pa, pb, pc, o, d, min_t, &max_t, &u, &v, w->mode)) {
                        index = i;
                    }
                }

                if (hit_p) {
                    break;
                }
            }

            n = stack[--sp];
        }

        if (w->any) {
            w->hits[r] = (uint32_t)hit_p;
        } else {
            uint32_t* out = w->hits + 4 * r;
            memcpy(out + 0, &max_t, 4);
            if (SENTINEL != index) {
                memcpy(out + 1, &u, 4);
                memcpy(out + 2, &v, 4);
                out[3] = w->slots[index];
            } else {
                out[1] = 0;
                out[2] = 0;
                out[3] = SENTINEL;
            }
        }

        if (w->counts) {
            w->counts[2 * r + 0] = num_nodes;
            w->counts[2 * r + 1] = num_tris;
        }
    }
}

static void* trace_thread(void* arg) {
    trace_range((otask*)arg);
    return NULL;
}

int32_t oracle_trace(onode const* nodes, uint32_t const* slots, float const* positions,
                     uint32_t const* indices, float const* rays, uint64_t num_rays, int32_t any,
                     int32_t rcp_mode, uint32_t* hits, uint32_t* counts, uint32_t num_threads) {
    if (0 == num_threads) {
        num_threads = 1;
    }
    if (num_threads > 256) {
        num_threads = 256;
    }

    otask     tasks[256];
    pthread_t threads[256];

    uint64_t const chunk = (num_rays + num_threads - 1) / num_threads;

    for (uint32_t t = 0; t < num_threads; ++t) {
        otask* w     = &tasks[t];
        w->nodes     = nodes;
        w->slots     = slots;
        w->positions = positions;
        w->indices   = indices;
        w->rays      = rays;
        w->begin     = t * chunk < num_rays ? t * chunk : num_rays;
        w->end       = (t + 1) * chunk < num_rays ? (t + 1) * chunk : num_rays;
        w->any       = any;
        w->mode      = rcp_mode;
        w->hits      = hits;
        w->counts    = counts;
        w->status    = 0;
    }

    if (1 == num_threads) {
        trace_range(&tasks[0]);
    } else {
        for (uint32_t t = 0; t < num_threads; ++t) {
            pthread_create(&threads[t], NULL, trace_thread, &tasks[t]);
        }
        for (uint32_t t = 0; t < num_threads; ++t) {
            pthread_join(threads[t], NULL);
        }
    }

    int32_t status = 0;
    for (uint32_t t = 0; t < num_threads; ++t) {
        if (tasks[t].status) {
            status = tasks[t].status;
        }
    }
    return status;
}

void oracle_brute_force(float const* positions, uint32_t const* indices, uint32_t num_triangles,
                        float const* rays, uint64_t num_rays, int32_t rcp_mode, uint32_t* hits) {
    for (uint64_t r = 0; r < num_rays; ++r) {
        float const* ray = rays + 8 * r;

        float const o[3]  = {ray[0], ray[1], ray[2]};
        float const d[3]  = {ray[4], ray[5], ray[6]};
        float const min_t = ray[3];
        float       max_t = ray[7];

        uint32_t index = SENTINEL;
        float    u = 0.f, v = 0.f;

        for (uint32_t t = 0; t < num_triangles; ++t) {
            float const* pa = positions + 3 * (size_t)indices[3 * t + 0];
            float const* pb = positions + 3 * (size_t)indices[3 * t + 1];
            float const* pc = positions + 3 * (size_t)indices[3 * t + 2];
            if (triangle(pa, pb, pc, o, d, min_t, &max_t, &u, &v, rcp_mode)) {
                index = t;
            }
        }

        uint32_t* out = hits + 4 * r;
        memcpy(out + 0, &max_t, 4);
        if (SENTINEL != index) {
            memcpy(out + 1, &u, 4);
            memcpy(out + 2, &v, 4);
        } else {
            out[1] = 0;
            out[2] = 0;
        }
        out[3] = index;
    }
}

/* offset_ray, core/scene/ray_offset.inl:18-29 */
void oracle_offset_ray(float const p[3], float const n[3], float out[3]) {
    float const origin      = 1.f / 32.f;
    float const float_scale = 1.f / 65536.f;
    float const int_scale   = 256.f;

    for (int i = 0; i < 3; ++i) {
        int32_t const of_i = (int32_t)(int_scale * n[i]);

        int32_t pi;
        memcpy(&pi, &p[i], 4);
        pi += (p[i] < 0.f) ? -of_i : of_i;

        float p_i;
        memcpy(&p_i, &pi, 4);

        float const ap = p[i] < 0.f ? -p[i] : p[i];
        out[i]         = ap < origin ? fmaf(float_scale, n[i], p[i]) : p_i;
    }
}
