// Thin C driver around the reference's own Substitute material sample
// (Disney diffuse + GGX, core/scene/material/substitute/substitute_sample.cpp:10-45,
// substitute_base_sample.inl, ggx/ggx.inl, disney/disney.inl), compiled from
// the reference sources where they lie (see ref.mk).  It sets up a
// substitute::Sample exactly as Material_base::set_sample does
// (substitute_base_material.inl:37-102) from plain parameters, and evaluates
// or samples it.  Test infrastructure only.

#include "base/math/vector3.inl"
#include "base/random/generator.inl"
#include "sampler/sampler.hpp"
#include "scene/material/bxdf.hpp"
#include "scene/material/fresnel/fresnel.inl"
#include "scene/material/ggx/ggx.inl"
#include "scene/material/material_sample.inl"
#include "scene/material/substitute/substitute_base_sample.inl"
#include "scene/material/substitute/substitute_sample.hpp"

#include <cstdint>
#include <cstring>

namespace {

// Returns preset numbers instead of a sequence: sample_1D then sample_2D.
class Fixed_sampler final : public sampler::Sampler {
  public:
    Fixed_sampler(float s1, float s2x, float s2y) : s1_(s1), s2_(s2x, s2y) {}

    float2 sample_2D(RNG& /*rng*/, uint32_t /*dimension*/) final {
        return s2_;
    }

    float sample_1D(RNG& /*rng*/, uint32_t /*dimension*/) final {
        return s1_;
    }

  private:
    float  s1_;
    float2 s2_;
};

using scene::material::substitute::Sample;

// params: color.rgb, roughness, metallic, ior
// frame:  t.xyz, b.xyz, n.xyz, geo_n.xyz
void setup(Sample& s, float const* params, float const* frame, float const* wo) {
    float3 const t(frame[0], frame[1], frame[2]);
    float3 const b(frame[3], frame[4], frame[5]);
    float3 const n(frame[6], frame[7], frame[8]);
    float3 const geo_n(frame[9], frame[10], frame[11]);

    s.layer_.set_tangent_frame(t, b, n);

    float const r     = scene::material::ggx::clamp_roughness(params[3]);
    float const alpha = r * r;

    s.set_common(geo_n, n, float3(wo[0], wo[1], wo[2]), alpha);

    float3 const color(params[0], params[1], params[2]);
    s.base_.set(color, scene::material::fresnel::schlick_f0(params[5], 1.f), params[4]);
}

}  // namespace

extern "C" {

// E(n_dot_wo, alpha) table the GGX energy compensation interpolates
// (ggx.inl:23, ggx_integral.inl), 32 x 32, row = alpha.
void sref_ggx_e_table(float* out) {
    std::memcpy(out, &scene::material::ggx::E[0][0], sizeof(scene::material::ggx::E));
}

// out per case: reflection.rgb (already times n_dot_wi), pdf
void sref_substitute_evaluate(uint32_t n, float const* params, float const* frames,
                              float const* wo, float const* wi, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        Sample s;
        setup(s, params + 6 * i, frames + 12 * i, wo + 3 * i);
        auto const r = s.evaluate(float3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]));
        out[4 * i + 0] = r.reflection[0];
        out[4 * i + 1] = r.reflection[1];
        out[4 * i + 2] = r.reflection[2];
        out[4 * i + 3] = r.pdf();
    }
}

// xi per case: lobe choice (1D), then the 2D sample.
// out per case: wi.xyz, reflection.rgb, pdf, type flags (as float bits)
void sref_substitute_sample(uint32_t n, float const* params, float const* frames,
                            float const* wo, float const* xi, float* out) {
    rnd::Generator rng(0, 0);
    for (uint32_t i = 0; i < n; ++i) {
        Sample s;
        setup(s, params + 6 * i, frames + 12 * i, wo + 3 * i);
        Fixed_sampler sampler(xi[3 * i], xi[3 * i + 1], xi[3 * i + 2]);
        scene::material::bxdf::Sample result;
        s.sample(sampler, rng, result);
        float* o = out + 8 * i;
        o[0]     = result.wi[0];
        o[1]     = result.wi[1];
        o[2]     = result.wi[2];
        o[3]     = result.reflection[0];
        o[4]     = result.reflection[1];
        o[5]     = result.reflection[2];
        o[6]     = result.pdf;
        uint32_t const type = result.pdf > 0.f ? uint32_t(result.type.values) : 0u;
        std::memcpy(&o[7], &type, 4);
    }
}

}  // extern "C"
