// Device Substitute material sample: Frostbite-normalized Disney diffuse plus
// height-correlated GGX with energy compensation, one-sample lobe choice.
//
//   substitute::Sample::evaluate / sample  core/scene/material/substitute/substitute_sample.cpp:10-45
//   Base_closure                           core/scene/material/substitute/substitute_base_sample.inl:17-238
//   disney::Iso                            core/scene/material/disney/disney.inl:21-77
//   ggx (isotropic), VNDF sampling, E      core/scene/material/ggx/ggx.inl:27-236,469-503
//   fresnel::schlick                       core/scene/material/fresnel/fresnel.inl:10-21
//   Layer, clamp_dot (Dot_min 1e-5)        core/scene/material/material_sample.inl:25-79
//
// Isotropic only (the configurations use no anisotropy).  sqrt, 1/sqrt and
// sin/cos are the correctly rounded / library versions, not the reference's
// SSE estimates.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mi::dev {

inline constexpr float Pi        = 3.14159265358979323846f;
inline constexpr float Pi_inv    = 0.318309886183790671538f;
inline constexpr float Dot_min   = 0.00001f;
inline constexpr float Min_rough = 0.01314f;
inline constexpr int   E_size    = 32;

inline constexpr uint32_t Type_diffuse  = 5;   // Reflection | Diffuse
inline constexpr uint32_t Type_glossy   = 9;   // Reflection | Glossy
inline constexpr uint32_t Type_specular = 17;  // Reflection | Specular

struct V3 {
    float x, y, z;
};

__device__ __forceinline__ V3 v3(float x, float y, float z) {
    return {x, y, z};
}
__device__ __forceinline__ V3 operator+(V3 a, V3 b) {
    return {a.x + b.x, a.y + b.y, a.z + b.z};
}
__device__ __forceinline__ V3 operator-(V3 a, V3 b) {
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
__device__ __forceinline__ V3 operator*(float s, V3 a) {
    return {s * a.x, s * a.y, s * a.z};
}
__device__ __forceinline__ V3 operator-(V3 a) {
    return {-a.x, -a.y, -a.z};
}
__device__ __forceinline__ float dot(V3 a, V3 b) {
    return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 normalize(V3 v) {
    float const s = 1.f / sqrtf(dot(v, v));
    return s * v;
}
__device__ __forceinline__ float clampf(float x, float lo, float hi) {
    return x < lo ? lo : (hi < x ? hi : x);
}
__device__ __forceinline__ float clamp_dot(V3 a, V3 b) {
    return clampf(dot(a, b), Dot_min, 1.f);
}
__device__ __forceinline__ float clamp_abs_dot(V3 a, V3 b) {
    return clampf(fabsf(dot(a, b)), Dot_min, 1.f);
}
__device__ __forceinline__ float pow5(float x) {
    float const x2 = x * x;
    return (x2 * x2) * x;
}
__device__ __forceinline__ float lerpf(float a, float b, float t) {
    float const u = 1.f - t;
    return u * a + t * b;
}

struct Layer {
    V3 t, b, n;

    __device__ V3 tangent_to_world(V3 v) const {
        return {v.x * t.x + v.y * b.x + v.z * n.x, v.x * t.y + v.y * b.y + v.z * n.y,
                v.x * t.z + v.y * b.z + v.z * n.z};
    }

    __device__ V3 world_to_tangent(V3 v) const {
        return {v.x * t.x + v.y * t.y + v.z * t.z, v.x * b.x + v.y * b.y + v.z * b.z,
                v.x * n.x + v.y * n.y + v.z * n.z};
    }
};

// orthonormal_basis, base/math/vector3.inl:523-534
__device__ __forceinline__ Layer frame_from_normal(V3 n) {
    float const sign = copysignf(1.f, n.z);  // copysign1, base/math/math.hpp:124-128
    float const c    = -1.f / (sign + n.z);
    float const d    = n.x * n.y * c;
    Layer       l;
    l.t = {1.f + sign * n.x * n.x * c, sign * d, -sign * n.x};
    l.b = {d, sign + n.y * n.y * c, -n.y};
    l.n = n;
    return l;
}

// sample_disk_concentric / sample_hemisphere_cosine, base/math/sampling.inl:11-32,77-82
__device__ __forceinline__ V3 hemisphere_cosine(float u, float v) {
    float const sx = 2.f * u - 1.f;
    float const sy = 2.f * v - 1.f;
    float       x  = 0.f;
    float       y  = 0.f;
    if (!(0.f == sx && 0.f == sy)) {
        float r, theta;
        if (fabsf(sx) > fabsf(sy)) {
            r     = sx;
            theta = (Pi / 4.f) * (sy / sx);
        } else {
            r     = sy;
            theta = (Pi / 2.f) - (Pi / 4.f) * (sx / sy);
        }
        x = cosf(theta) * r;
        y = sinf(theta) * r;
    }
    return {x, y, sqrtf(fmaxf(0.f, 1.f - x * x - y * y))};
}

struct Bxdf_sample {
    V3       wi;
    V3       h;
    float    refl[3];
    float    pdf;
    float    h_dot_wi;
    uint32_t type;
};

// The energy table E(n.wo, alpha) lives in device memory, filled once by
// make_energy_table (bsdf_table.hip).
struct Material_closure {
    Layer        layer;
    V3           geo_n;
    V3           wo;
    float        alpha;
    float        albedo[3];
    float        f0[3];
    float        metallic;
    float const* e_table;  // E_size * E_size, row = alpha

    // Material_base::set_sample + Base_closure::set
    // (substitute_base_material.inl:37-102, substitute_base_sample.inl:17-24)
    __device__ void set(Layer const& l, V3 geo, V3 w, float const color[3], float roughness,
                        float metal, float ior, float ior_outside, float const* table) {
        layer           = l;
        geo_n           = geo;
        wo              = w;
        float const r   = roughness > Min_rough ? roughness : Min_rough;
        alpha           = r * r;
        float const t   = (ior - ior_outside) / (ior + ior_outside);
        float const f0s = t * t;
        for (int c = 0; c < 3; ++c) {
            albedo[c] = (1.f - metal) * color[c];
            f0[c]     = lerpf(f0s, color[c], metal);
        }
        metallic = metal;
        e_table  = table;
    }

    __device__ bool same_hemisphere(V3 v) const { return dot(geo_n, v) > 0.f; }

    // Interpolated_function_2D_N::operator()(n_dot_wo, alpha)
    __device__ float e_lookup(float x, float y) const {
        x                   = fminf(x, 1.f);
        y                   = fminf(y, 1.f);
        float const    ox   = x * float(E_size - 1);
        float const    oy   = y * float(E_size - 1);
        uint32_t const ix   = uint32_t(ox);
        uint32_t const iy   = uint32_t(oy);
        float const    tx   = ox - float(ix);
        float const    ty   = oy - float(iy);
        uint32_t const col1 = ix + 1 < E_size - 1 ? ix + 1 : E_size - 1;
        uint32_t const row0 = iy * E_size;
        uint32_t const row1 = (iy + 1 < E_size - 1 ? iy + 1 : E_size - 1) * E_size;
        float const    sx   = 1.f - tx;
        float const    sy   = 1.f - ty;
        return sy * (sx * e_table[ix + row0] + tx * e_table[col1 + row0]) +
               ty * (sx * e_table[ix + row1] + tx * e_table[col1 + row1]);
    }

    __device__ void ilm(float n_dot_wo, float out[3]) const {
        float const e = e_lookup(n_dot_wo, alpha);
        for (int c = 0; c < 3; ++c) {
            out[c] = 1.f + (metallic / e - 1.f) * f0[c];
        }
    }

    __device__ void disney(float h_dot_wi, float n_dot_wi, float n_dot_wo, float out[3]) const {
        float const energy_bias   = lerpf(0.f, 0.5f, alpha);
        float const energy_factor = lerpf(1.f, 1.f / 1.53f, alpha);
        float const f_d90         = energy_bias + (2.f * alpha) * (h_dot_wi * h_dot_wi);
        float const fmo           = f_d90 - 1.f;
        float const a             = 1.f + fmo * pow5(1.f - n_dot_wi);
        float const b             = 1.f + fmo * pow5(1.f - n_dot_wo);
        for (int c = 0; c < 3; ++c) {
            out[c] = (a * b * energy_factor) * (Pi_inv * albedo[c]);
        }
    }

    __device__ void ggx(float n_dot_wi, float n_dot_wo, float wo_dot_h, float n_dot_h,
                        float out[3], float& pdf) const {
        float const a2   = alpha * alpha;
        float const dd   = (n_dot_h * n_dot_h) * (a2 - 1.f) + 1.f;
        float const d    = a2 / (Pi * dd * dd);
        float const t_wi = sqrtf(a2 + (1.f - a2) * (n_dot_wi * n_dot_wi));
        float const t_wo = sqrtf(a2 + (1.f - a2) * (n_dot_wo * n_dot_wo));
        float const vis  = 0.5f / (n_dot_wi * t_wo + n_dot_wo * t_wi);
        float const p5   = pow5(1.f - wo_dot_h);
        for (int c = 0; c < 3; ++c) {
            out[c] = d * vis * (f0[c] + p5 * (1.f - f0[c]));
        }
        pdf = (0.5f * d) / (t_wo + n_dot_wo);
    }

    // Aniso::sample with equal alphas (VNDF), ggx.inl:469-503
    __device__ V3 sample_h(float xi0, float xi1, float& n_dot_h) const {
        V3 const lwo = layer.world_to_tangent(wo);
        V3 const v   = normalize(v3(alpha * lwo.x, alpha * lwo.y, lwo.z));

        V3 const cross_v_z = normalize(v3(v.y, -v.x, 0.f));
        V3 const t1        = (v.z < 0.9999f) ? cross_v_z : v3(1.f, 0.f, 0.f);
        V3 const t2        = v3(t1.y * v.z, -t1.x * v.z, t1.x * v.y - t1.y * v.x);

        float const a   = 1.f / (1.f + v.z);
        float const r   = sqrtf(xi0);
        float const phi = (xi1 < a) ? (xi1 / a * Pi) : (Pi + (xi1 - a) / (1.f - a) * Pi);

        float const p1 = r * cosf(phi);
        float const p2 = r * sinf(phi) * ((xi1 < a) ? 1.f : v.z);

        V3 m = (p1 * t1 + p2 * t2) + sqrtf(fmaxf(1.f - p1 * p1 - p2 * p2, 0.f)) * v;
        m    = normalize(v3(alpha * m.x, alpha * m.y, fmaxf(m.z, 0.f)));

        n_dot_h = clampf(m.z, 0.f, 1.f);
        return layer.tangent_to_world(m);
    }

    // Iso::reflect: returns n_dot_wi
    __device__ float ggx_reflect(float n_dot_wo, float xi0, float xi1, Bxdf_sample& s) const {
        float       n_dot_h;
        V3 const    h        = sample_h(xi0, xi1, n_dot_h);
        float const wo_dot_h = clamp_dot(wo, h);
        V3 const    wi       = normalize((2.f * wo_dot_h) * h - wo);
        float const n_dot_wi = clamp_dot(layer.n, wi);
        ggx(n_dot_wi, n_dot_wo, wo_dot_h, n_dot_h, s.refl, s.pdf);
        s.wi       = wi;
        s.h        = h;
        s.h_dot_wi = wo_dot_h;
        s.type     = alpha <= Min_rough * Min_rough ? Type_specular : Type_glossy;
        return n_dot_wi;
    }

    // Sample::evaluate: out = reflection (times n.wi), pdf
    __device__ void evaluate(V3 wi, float out[4]) const {
        if (!same_hemisphere(wo)) {
            out[0] = out[1] = out[2] = out[3] = 0.f;
            return;
        }
        V3 const    h        = normalize(wo + wi);
        float const wo_dot_h = clamp_dot(wo, h);
        float const n_dot_wi = clamp_dot(layer.n, wi);
        float const n_dot_wo = clamp_abs_dot(layer.n, wo);
        float const n_dot_h  = clampf(dot(layer.n, h), 0.f, 1.f);

        float g[3], g_pdf, e[3];
        ggx(n_dot_wi, n_dot_wo, wo_dot_h, n_dot_h, g, g_pdf);
        ilm(n_dot_wo, e);

        if (1.f == metallic) {
            for (int c = 0; c < 3; ++c) {
                out[c] = n_dot_wi * (g[c] * e[c]);
            }
            out[3] = g_pdf;
            return;
        }
        float d[3];
        disney(wo_dot_h, n_dot_wi, n_dot_wo, d);
        for (int c = 0; c < 3; ++c) {
            out[c] = n_dot_wi * (d[c] + g[c] * e[c]);
        }
        out[3] = 0.5f * (n_dot_wi * Pi_inv + g_pdf);
    }

    // Sample::sample with the lobe choice xi_lobe and the 2D sample (xi0, xi1)
    __device__ void sample(float xi_lobe, float xi0, float xi1, Bxdf_sample& s) const {
        s.pdf  = 0.f;
        s.type = 0;
        if (!same_hemisphere(wo)) {
            return;
        }
        float const n_dot_wo = clamp_abs_dot(layer.n, wo);
        float       e[3];
        ilm(n_dot_wo, e);

        if (1.f == metallic) {
            float const n_dot_wi = ggx_reflect(n_dot_wo, xi0, xi1, s);
            for (int c = 0; c < 3; ++c) {
                s.refl[c] *= n_dot_wi * e[c];
            }
            return;
        }

        if (xi_lobe < 0.5f) {
            V3 const    is       = hemisphere_cosine(xi0, xi1);
            V3 const    wi       = normalize(layer.tangent_to_world(is));
            V3 const    h        = normalize(wo + wi);
            float const h_dot_wi = clamp_dot(h, wi);
            float const n_dot_wi = clamp_dot(layer.n, wi);
            float       d[3];
            disney(h_dot_wi, n_dot_wi, n_dot_wo, d);
            float const n_dot_h = clampf(dot(layer.n, h), 0.f, 1.f);
            float       g[3], g_pdf;
            ggx(n_dot_wi, n_dot_wo, h_dot_wi, n_dot_h, g, g_pdf);
            for (int c = 0; c < 3; ++c) {
                s.refl[c] = n_dot_wi * (d[c] + g[c] * e[c]);
            }
            s.wi       = wi;
            s.h        = h;
            s.h_dot_wi = h_dot_wi;
            s.pdf      = 0.5f * (n_dot_wi * Pi_inv + g_pdf);
            s.type     = Type_diffuse;
        } else {
            float const n_dot_wi = ggx_reflect(n_dot_wo, xi0, xi1, s);
            float       d[3];
            disney(s.h_dot_wi, n_dot_wi, n_dot_wo, d);
            for (int c = 0; c < 3; ++c) {
                s.refl[c] = n_dot_wi * (s.refl[c] * e[c] + d[c]);
            }
            s.pdf = 0.5f * (s.pdf + n_dot_wi * Pi_inv);
        }
    }
};

}  // namespace mi::dev
