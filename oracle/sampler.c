/* ORACLE (test infrastructure only): sprout's camera samples and perspective
 * camera ray, restated on the CPU to check the GPU's raygen.  Nothing in the
 * product links this file.
 *
 *   rnd::Generator (PCG32)            base/random/generator.inl:16-49
 *   math::golden_ratio (1-D, 2-D)     base/math/sample_distribution.inl:68-147
 *   rnd::biased_shuffle               base/random/shuffle.hpp:12-32
 *   Golden_ratio::generate_2D / 1D    core/sampler/sampler_golden_ratio.cpp:51-73
 *   Sampler::camera_sample            core/sampler/sampler.cpp:11-17
 *   per-pixel generator start         core/rendering/rendering_worker.cpp:97-104
 *   Perspective::on_update            core/scene/camera/camera_perspective.cpp:204-226
 *   Perspective::generate_ray         core/scene/camera/camera_perspective.cpp:47-80
 *   Composed_transformation::object_to_world_vector
 *                                     core/scene/composed_transformation.inl:84-93
 *
 * Floating-point order follows the reference as compiled (g++ -O3
 * -march=native, which contracts products into fused multiply-adds; the x86
 * assembly of generate_ray built here from the reference source shows the
 * order written out below with fmaf).  The reference's normalize takes the
 * SSE rsqrt estimate and one Newton step; here the estimate is the correctly
 * rounded 1/sqrt, followed by the same Newton step (the PINNED mode of the
 * traversal's reciprocal).
 */

#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define GOLDEN_STEP 0.618033988749894f

/* math::golden_ratio over floats: start at r, add the golden step, wrap at 1 */
static void golden_ratio_1d(float* s, uint32_t n, float r) {
    float x = r;
    for (uint32_t i = 0; i < n; ++i) {
        s[i] = x;
        x += GOLDEN_STEP;
        if (x >= 1.f) {
            x -= 1.f;
        }
    }
}

/* math::golden_ratio over float2 (interleaved x, y): the second coordinates
 * of the r[0] sequence, read in the Fibonacci order that starts at their
 * minimum, become the first coordinates; the r[1] sequence the second. */
static void golden_ratio_2d(float* s, uint32_t n, float r0, float r1) {
    float    x     = r0;
    float    least = x;
    uint32_t idx   = 0;
    for (uint32_t i = 0; i < n; ++i) {
        s[2 * i + 1] = x;
        if (x < least) {
            least = x;
            idx   = i;
        }
        x += GOLDEN_STEP;
        if (x >= 1.f) {
            x -= 1.f;
        }
    }

    /* the smallest Fibonacci pair f + fp >= n and the parity of its index */
    uint32_t f = 1, fp = 1, parity = 0;
    while (f + fp < n) {
        uint32_t const t = f;
        f += fp;
        fp = t;
        ++parity;
    }
    uint32_t const inc = (parity & 1u) ? f : fp;
    uint32_t const dec = (parity & 1u) ? fp : f;

    s[0] = s[2 * idx + 1];
    for (uint32_t i = 1; i < n; ++i) {
        if (idx < dec) {
            idx += inc;
            if (idx >= n) {
                idx -= dec;
            }
        } else {
            idx -= dec;
        }
        s[2 * i] = s[2 * idx + 1];
    }

    float y = r1;
    for (uint32_t i = 0; i < n; ++i) {
        s[2 * i + 1] = y;
        y += GOLDEN_STEP;
        if (y >= 1.f) {
            y -= 1.f;
        }
    }
}

/* rnd::biased_shuffle: Fisher-Yates from the back, the partner index the
 * high word of random_uint * (i + 1) (Lemire's multiply, no rejection) */
static void shuffle_elements(float* s, uint32_t n, uint32_t width, oracle_pcg* g) {
    if (n < 2) {
        return;
    }
    for (uint64_t i = (uint64_t)n - 1; i > 0; --i) {
        uint64_t const r = (uint64_t)oracle_pcg_uint(g);
        uint64_t const o = (r * (i + 1)) >> 32;
        for (uint32_t c = 0; c < width; ++c) {
            float const t        = s[width * i + c];
            s[width * i + c]     = s[width * o + c];
            s[width * o + c]     = t;
        }
    }
}

uint64_t oracle_pixel_sequence(int32_t x, int32_t y, uint32_t grid_width, uint32_t grid_height,
                               int32_t border, uint32_t iteration) {
    uint64_t const o0 = (uint64_t)iteration * ((uint64_t)grid_width * grid_height);
    uint64_t const o1 = (uint64_t)(int64_t)((y + border) * (int32_t)grid_width) + o0;
    return o1 + (uint64_t)(int64_t)(x + border);
}

int32_t oracle_golden_ratio_camera_samples(uint32_t spp, uint64_t sequence, float* image,
                                           float* lens, float* time) {
    if (0 == spp || !image) {
        return ORACLE_ERR_ARGS;
    }
    oracle_pcg g;
    oracle_pcg_start(&g, 0, sequence);

    /* sample_2D(rng, 0) of the first sample: the image dimension.  The start
     * pair is float2(random_float(), random_float()); g++ evaluates those
     * constructor arguments right to left, so the first draw is r[1]. */
    float const b = oracle_pcg_float(&g);
    float const a = oracle_pcg_float(&g);
    golden_ratio_2d(image, spp, a, b);
    shuffle_elements(image, spp, 2, &g);

    /* sample_2D(rng, 1): the lens dimension */
    float* l = lens ? lens : (float*)malloc(sizeof(float) * 2 * spp);
    float const d = oracle_pcg_float(&g);
    float const c = oracle_pcg_float(&g);
    golden_ratio_2d(l, spp, c, d);
    shuffle_elements(l, spp, 2, &g);
    if (!lens) {
        free(l);
    }

    /* sample_1D(rng, 0): the time dimension */
    float* t = time ? time : (float*)malloc(sizeof(float) * spp);
    golden_ratio_1d(t, spp, oracle_pcg_float(&g));
    shuffle_elements(t, spp, 1, &g);
    if (!time) {
        free(t);
    }
    return 0;
}

void oracle_perspective_setup(uint32_t width, uint32_t height, float fov_deg,
                              oracle_perspective* cam) {
    float const pi_over_180 = 3.14159265358979323846f / 180.f;
    float const fov         = fov_deg * pi_over_180;
    float const fw          = (float)width;
    float const fh          = (float)height;
    float const ratio       = fh / fw;
    float const z           = 1.f / tanf(0.5f * fov);
    /* left_top (-1, ratio, z), right_top (1, ratio, z), left_bottom (-1, -ratio, z) */
    cam->left_top[0] = -1.f;
    cam->left_top[1] = ratio;
    cam->left_top[2] = z;
    cam->d_x[0]      = (1.f - -1.f) / fw;
    cam->d_x[1]      = (ratio - ratio) / fw;
    cam->d_x[2]      = (z - z) / fw;
    cam->d_y[0]      = (-1.f - -1.f) / fh;
    cam->d_y[1]      = (-ratio - ratio) / fh;
    cam->d_y[2]      = (z - z) / fh;
}

void oracle_perspective_ray(oracle_perspective const* cam, float const position[3],
                            float const rotation[9], int32_t x, int32_t y, float u, float v,
                            float origin[3], float direction[3]) {
    float const c0 = (float)x + u;
    float const c1 = (float)y + v;
    float       d[3];
    for (int k = 0; k < 3; ++k) {
        d[k] = fmaf(c1, cam->d_y[k], fmaf(c0, cam->d_x[k], cam->left_top[k]));
    }
    /* normalize: dot as fma(z, z, fma(x, x, y y)), then the Newton step on the
     * estimate 1/sqrt */
    float const dot = fmaf(d[2], d[2], fmaf(d[0], d[0], d[1] * d[1]));
    float const r   = 1.f / sqrtf(dot);
    float const s   = (0.5f * r) * (3.f - (dot * r) * r);
    float const n[3] = {d[0] * s, d[1] * s, d[2] * s};

    /* object_to_world_vector with unit scale: rows r0 (right), r1 (up), r2
     * (view direction); x and y as fma(n.z, r2, fma(n.y, r1, n.x r0)), z as
     * fma(n.z, r2, fma(n.x, r0, n.y r1)) */
    float const* r0 = rotation;
    float const* r1 = rotation + 3;
    float const* r2 = rotation + 6;
    direction[0]    = fmaf(n[2], r2[0], fmaf(n[1], r1[0], n[0] * r0[0]));
    direction[1]    = fmaf(n[2], r2[1], fmaf(n[1], r1[1], n[0] * r0[1]));
    direction[2]    = fmaf(n[2], r2[2], fmaf(n[0], r0[2], n[1] * r1[2]));
    origin[0]       = position[0];
    origin[1]       = position[1];
    origin[2]       = position[2];
}
