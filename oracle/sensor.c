/*
 * ORACLE (test infrastructure only; never linked into the product).
 *
 * CPU restatement of sprout's sensor filters:
 *   Interpolated_function_1D_N<float, 30>  base/math/interpolated_function_1d.inl:124-178
 *   lerp                                    base/math/math.hpp:56-60
 *   Gaussian(radius, alpha)                 core/rendering/sensor/filter/sensor_gaussian.cpp
 *   Blackman(radius)                        core/rendering/sensor/filter/sensor_blackman.cpp
 *   Mitchell(radius, b, c)                  core/rendering/sensor/filter/sensor_mitchell.cpp
 *   defaults (radius 2, alpha 1.8, b = c = 1/3)  core/take/take_loader.cpp:331-376
 * and of the clamps (core/rendering/sensor/clamp.inl:9-35).  The splat of
 * Filtered_1p0 / Filtered_2p0::add_sample (core/rendering/sensor/filtered.inl:
 * 121-300) is restated in oracle.py over the samples pt.c returns.
 * Pinned against the reference's own filter code built here (oracle/ref.mk).
 */

#include "oracle.h"

#include <math.h>
#include <string.h>

#define TABLE 30

static float table_eval(oracle_filter const* f, float x) {
    x = fminf(x, f->range_end);
    float const    o      = x * f->inverse_interval;
    uint32_t const offset = (uint32_t)o;
    float const    t      = o - (float)offset;
    uint32_t const next   = offset + 1 < TABLE - 1 ? offset + 1 : TABLE - 1;
    float const    u      = 1.f - t;
    return u * f->table[offset] + t * f->table[next];
}

static float mitchell(float b, float c, float x) {
    x              = fabsf(2.f * x);
    float const xx = x * x;
    if (x > 1.f) {
        return ((((-b - 6.f * c) * xx) * x + (6.f * b + 30.f * c) * xx) + (-12.f * b - 48.f * c) * x +
                (8.f * b + 24.f * c)) *
               (1.f / 6.f);
    }
    return ((((12.f - 9.f * b) - 6.f * c) * xx) * x + ((-18.f + 12.f * b) + 6.f * c) * xx +
            (6.f - 2.f * b)) *
           (1.f / 6.f);
}

float oracle_filter_eval(oracle_filter const* f, float d) {
    switch (f->kind) {
        case ORACLE_FILTER_GAUSSIAN:
            return table_eval(f, d * d);
        case ORACLE_FILTER_BLACKMAN:
            return table_eval(f, fabsf(d));
        case ORACLE_FILTER_MITCHELL:
            return mitchell(f->b, f->c, d * f->radius_inv);
        default:
            return fabsf(d) <= 0.5f ? 1.f : 0.f;
    }
}

/* the table of f over [0, end], sample i at the running sum of the interval */
static void fill(oracle_filter* f, float end) {
    float const range    = end;
    float const interval = range / (float)(TABLE - 1);
    f->range_end         = end;
    f->inverse_interval  = 1.f / interval;
    float s              = 0.f;
    for (int i = 0; i < TABLE; ++i, s += interval) {
        if (ORACLE_FILTER_GAUSSIAN == f->kind) {
            float const e = expf(-f->alpha * (f->radius * f->radius));
            f->table[i]   = fmaxf(0.f, expf(-f->alpha * s) - e);
        } else {
            float const a0 = 0.35875f, a1 = 0.48829f, a2 = 0.14128f, a3 = 0.01168f;
            float const x  = (3.14159265358979323846f * (s + f->radius)) / f->radius;
            f->table[i]    = ((a0 - a1 * cosf(x)) + a2 * cosf(2.f * x)) - a3 * cosf(3.f * x);
        }
    }
}

/* midpoint rule over [0, radius] with 64 samples, doubled */
static float integral(oracle_filter const* f) {
    float const interval = f->radius / 64.f;
    float       s        = 0.5f * interval;
    float       sum      = 0.f;
    for (int i = 0; i < 64; ++i, s += interval) {
        sum += oracle_filter_eval(f, s) * interval;
    }
    return sum + sum;
}

void oracle_filter_init(oracle_filter* f, int32_t kind, float radius, float alpha, float b,
                        float c) {
    memset(f, 0, sizeof(*f));
    f->kind       = kind;
    f->radius     = radius;
    f->alpha      = alpha;
    f->b          = b;
    f->c          = c;
    f->radius_inv = 1.f / radius;
    if (ORACLE_FILTER_GAUSSIAN == kind || ORACLE_FILTER_BLACKMAN == kind) {
        fill(f, ORACLE_FILTER_GAUSSIAN == kind ? radius * radius : radius);
        float const s = 1.f / integral(f);
        for (int i = 0; i < TABLE; ++i) {
            f->table[i] *= s;
        }
    }
}

void oracle_filter_eval_many(oracle_filter const* f, float const* d, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        out[i] = oracle_filter_eval(f, d[i]);
    }
}
