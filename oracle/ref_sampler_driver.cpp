// C entry points over the reference's random generator, Golden_ratio camera
// sampler and sRGB -> ACEScg (AP1) colour mapping, built by ref.mk into
// oracle/_ref.  Test infrastructure only.

#include "base/random/generator.inl"
#include "base/spectrum/aces.hpp"
#include "sampler/camera_sample.hpp"
#include "sampler/sampler_golden_ratio.hpp"

#include <cstdint>

extern "C" {

// n draws of rnd::Generator(seed, sequence).random_uint()
void sref_pcg_uint(uint64_t seed, uint64_t sequence, uint32_t n, uint32_t* out) {
    rnd::Generator rng(seed, sequence);
    for (uint32_t i = 0; i < n; ++i) {
        out[i] = rng.random_uint();
    }
}

// n draws of rnd::Generator(seed, sequence).random_float()
void sref_pcg_float(uint64_t seed, uint64_t sequence, uint32_t n, float* out) {
    rnd::Generator rng(seed, sequence);
    for (uint32_t i = 0; i < n; ++i) {
        out[i] = rng.random_float();
    }
}

// The camera samples of one pixel as Worker::render draws them: the worker's
// generator started at (0, sequence), a Golden_ratio sampler of 2 2-D and 1
// 1-D dimensions for spp samples, start_pixel, then spp camera_sample calls.
// image, lens: 2 spp floats; time: spp floats.
void sref_golden_ratio_camera_samples(uint32_t spp, uint64_t sequence, int32_t x, int32_t y,
                                      float* image, float* lens, float* time) {
    rnd::Generator rng;
    rng.start(0, sequence);
    sampler::Golden_ratio sampler(2, 1, spp);
    sampler.start_pixel(rng);
    for (uint32_t i = 0; i < spp; ++i) {
        sampler::Camera_sample const s = sampler.camera_sample(rng, int2(x, y));
        image[2 * i + 0]               = s.pixel_uv[0];
        image[2 * i + 1]               = s.pixel_uv[1];
        lens[2 * i + 0]                = s.lens_uv[0];
        lens[2 * i + 1]                = s.lens_uv[1];
        time[i]                        = s.time;
    }
}

// spectrum::sRGB_to_AP1 of n colours (3 floats each)
void sref_srgb_to_ap1(float const* in, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        float3 const c = spectrum::sRGB_to_AP1(float3(in[3 * i], in[3 * i + 1], in[3 * i + 2]));
        out[3 * i + 0] = c[0];
        out[3 * i + 1] = c[1];
        out[3 * i + 2] = c[2];
    }
}
}
