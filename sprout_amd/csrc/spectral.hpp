// Hero-wavelength spectral transport (Wilkie, Nawaz, Droske, Weidlich and
// Hanika 2014, "Hero Wavelength Spectral Sampling"): every path carries four
// wavelengths over 380-720 nm, a uniformly drawn hero and three more at
// quarter-range offsets (rotated into the range), and its throughput per
// wavelength.  sprout renders in RGB only; this mode has no counterpart in it.
//
// RGB inputs (material colours, emission, the medium's coefficients) become
// spectra through a partition of unity of three piecewise-linear bands:
//
//   blue  1 below 450 nm, falling to 0 at 510 nm
//   red   0 below 560 nm, rising to 1 at 610 nm
//   green 1 - blue - red
//
// so x(lambda) = r x_r + g x_g + b x_b, and a grey input stays constant over
// wavelength.  The film reads RGB back through the dual basis: with the Gram
// matrix M_ij = E_lambda[b_i b_j] (lambda uniform over the range), the
// response of channel i is sum_j (M^-1)_ij b_j(lambda), so a spectrum in the
// span of the bands reads back as its own RGB, and the four-wavelength
// estimate rgb_i = 1/4 sum_k response_i(lambda_k) L(lambda_k) is unbiased.
// A scene whose surfaces and media are grey renders the RGB image in
// expectation; chromatic scattering differs from RGB as spectral transport
// does.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mi::spectral {

inline constexpr int   Num_wavelengths = 4;
inline constexpr float Lambda_min      = 380.f;
inline constexpr float Lambda_range    = 340.f;

// PCG hash (Jarzynski and Olano 2020, "Hash Functions for GPU Rendering",
// JCGT 9(3)): one round of the PCG output permutation over a 32-bit state.
__host__ __device__ inline uint32_t pcg_hash(uint32_t v) {
    uint32_t const state = v * 747796405u + 2891336453u;
    uint32_t const word  = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

// The hero wavelength's uniform number in [0, 1) of the path with generator
// sequence `sequence` under `seed`.  A hash of its own, so that the path's
// generator draws the same numbers as in RGB mode.
__host__ __device__ inline float hero_u(uint64_t seed, uint64_t sequence) {
    uint32_t h = pcg_hash(uint32_t(seed) ^ pcg_hash(uint32_t(seed >> 32)));
    h          = pcg_hash(h ^ uint32_t(sequence));
    h          = pcg_hash(h ^ uint32_t(sequence >> 32));
    return float(h >> 8) * 0x1p-24f;
}

// Wavelength k of a path with hero number u.
__host__ __device__ inline float wavelength(float u, int k) {
    float uk = u + 0.25f * float(k);
    if (uk >= 1.f) {
        uk -= 1.f;
    }
    return __builtin_fmaf(Lambda_range, uk, Lambda_min);
}

// Band weights (red, green, blue) at lambda.
__host__ __device__ inline void bands(float lambda, float b[3]) {
    float blue = (510.f - lambda) / 60.f;
    float red  = (lambda - 560.f) / 50.f;
    blue       = blue < 0.f ? 0.f : (blue > 1.f ? 1.f : blue);
    red        = red < 0.f ? 0.f : (red > 1.f ? 1.f : red);
    b[0]       = red;
    b[1]       = (1.f - blue) - red;
    b[2]       = blue;
}

// x(lambda) of an rgb value under band weights b.
__host__ __device__ inline float upsample(float const b[3], float const x[3]) {
    return __builtin_fmaf(b[2], x[2], __builtin_fmaf(b[1], x[1], b[0] * x[0]));
}

// The band weights of a path's four wavelengths.
struct Wavelengths {
    float b[Num_wavelengths][3];

    __host__ __device__ void set(float u) {
        for (int k = 0; k < Num_wavelengths; ++k) {
            bands(wavelength(u, k), b[k]);
        }
    }

    // x (rgb) at the four wavelengths
    __host__ __device__ void upsample4(float const x[3], float out[Num_wavelengths]) const {
        for (int k = 0; k < Num_wavelengths; ++k) {
            out[k] = upsample(b[k], x);
        }
    }

    // The film's rgb of values at the four wavelengths: minv is M^-1 (row
    // major); response_i(lambda_k) / 4 weighs wavelength k.
    __host__ __device__ void to_rgb(float const minv[9], float const v[Num_wavelengths],
                                    float rgb[3]) const {
        for (int i = 0; i < 3; ++i) {
            float sum = 0.f;
            for (int k = 0; k < Num_wavelengths; ++k) {
                float const r = 0.25f * __builtin_fmaf(minv[3 * i + 2], b[k][2],
                                                       __builtin_fmaf(minv[3 * i + 1], b[k][1],
                                                                      minv[3 * i + 0] * b[k][0]));
                sum = __builtin_fmaf(r, v[k], sum);
            }
            rgb[i] = sum;
        }
    }
};

// M^-1 of the bands' Gram matrix over the uniform wavelength measure (host):
// each product of two bands is quadratic between the knots 380, 450, 510,
// 560, 610 and 720 nm, so Simpson's rule per piece is exact.
void response_matrix(float minv[9]);

}  // namespace mi::spectral
