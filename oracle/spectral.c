/*
 * ORACLE (test infrastructure only; never linked into the product).
 * Hero-wavelength transport as the path tracer's spectral mode defines it
 * (sprout renders RGB only; the mode has no counterpart in the reference):
 *
 *   - four wavelengths per path over [380, 720) nm: a hero from
 *     u = hash(seed, generator sequence) and three more at u + k / 4 (mod 1)
 *     (Wilkie et al. 2014, "Hero Wavelength Spectral Sampling");
 *   - the hash is the PCG hash of Jarzynski and Olano 2020 ("Hash Functions
 *     for GPU Rendering", JCGT 9(3)), chained over the 32-bit halves;
 *   - rgb values become spectra through three piecewise-linear bands that
 *     sum to one (blue 1 below 450 nm falling to 0 at 510 nm, red 0 below
 *     560 nm rising to 1 at 610 nm, green the rest);
 *   - the film reads rgb through the bands' dual basis, the inverse of
 *     their Gram matrix over the uniform wavelength measure, each of the
 *     four wavelengths weighted 1/4.
 */

#include "oracle.h"

#include <math.h>

static uint32_t pcg_hash(uint32_t v) {
    uint32_t const state = v * 747796405u + 2891336453u;
    uint32_t const word  = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

float oracle_spectral_hero(uint64_t seed, uint64_t sequence) {
    uint32_t h = pcg_hash((uint32_t)seed ^ pcg_hash((uint32_t)(seed >> 32)));
    h          = pcg_hash(h ^ (uint32_t)sequence);
    h          = pcg_hash(h ^ (uint32_t)(sequence >> 32));
    return (float)(h >> 8) * 0x1p-24f;
}

float oracle_spectral_wavelength(float u, int32_t k) {
    float uk = u + 0.25f * (float)k;
    if (uk >= 1.f) {
        uk -= 1.f;
    }
    return fmaf(340.f, uk, 380.f);
}

void oracle_spectral_bands(float lambda, float b[3]) {
    float blue = (510.f - lambda) / 60.f;
    float red  = (lambda - 560.f) / 50.f;
    blue       = blue < 0.f ? 0.f : (blue > 1.f ? 1.f : blue);
    red        = red < 0.f ? 0.f : (red > 1.f ? 1.f : red);
    b[0]       = red;
    b[1]       = (1.f - blue) - red;
    b[2]       = blue;
}

float oracle_spectral_upsample(float const b[3], float const x[3]) {
    return fmaf(b[2], x[2], fmaf(b[1], x[1], b[0] * x[0]));
}

static void bands_d(double lambda, double b[3]) {
    double blue = (510.0 - lambda) / 60.0;
    double red  = (lambda - 560.0) / 50.0;
    blue        = blue < 0.0 ? 0.0 : (blue > 1.0 ? 1.0 : blue);
    red         = red < 0.0 ? 0.0 : (red > 1.0 ? 1.0 : red);
    b[0]        = red;
    b[1]        = (1.0 - blue) - red;
    b[2]        = blue;
}

void oracle_spectral_response(float minv[9]) {
    /* Gram matrix: products of two bands are quadratic between the knots,
     * so Simpson's rule per piece is exact */
    static double const knots[6] = {380.0, 450.0, 510.0, 560.0, 610.0, 720.0};
    double              m[9]     = {0};
    for (int p = 0; p < 5; ++p) {
        double const a = knots[p], c = knots[p + 1];
        double       ba[3], bm[3], bc[3];
        bands_d(a, ba);
        bands_d(0.5 * (a + c), bm);
        bands_d(c, bc);
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) {
                m[3 * i + j] += (c - a) / 6.0 * (ba[i] * ba[j] + 4.0 * bm[i] * bm[j] + bc[i] * bc[j]);
            }
        }
    }
    for (int k = 0; k < 9; ++k) {
        m[k] /= 340.0;
    }
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
        minv[k] = (float)inv[k];
    }
}

void oracle_spectral_to_rgb(float const minv[9], float const b[4][3], float const v[4],
                            float rgb[3]) {
    for (int i = 0; i < 3; ++i) {
        float sum = 0.f;
        for (int k = 0; k < 4; ++k) {
            float const r =
                0.25f * fmaf(minv[3 * i + 2], b[k][2], fmaf(minv[3 * i + 1], b[k][1], minv[3 * i + 0] * b[k][0]));
            sum = fmaf(r, v[k], sum);
        }
        rgb[i] = sum;
    }
}
