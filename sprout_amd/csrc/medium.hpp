// Heterogeneous participating medium: a density grid in an axis-aligned
// box, its majorant gridtree, and RGB delta / ratio tracking through it.
//
//   Grid::collision_coefficients, density  core/scene/material/volumetric/volumetric_grid.cpp:35-70
//   Linear_3D sampling, clamp addressing   core/image/texture/texture_sampler.cpp:175-188, 256-282,
//                                          core/image/texture/address_mode.hpp
//   Gridtree::intersect                    core/scene/material/volumetric/volumetric_octree.cpp:53-129
//   Octree_builder (host)                  core/scene/material/volumetric/volumetric_octree_builder.cpp:13-296
//   Tracking::tracking (scatter)           core/rendering/integrator/volume/tracking.cpp:328-375
//   tracking_transmitted, residual ratio   tracking.cpp:32-129
//   Tracking_multi::integrate (cell walk)  core/rendering/integrator/volume/tracking_multi.cpp:44-183
//   texture_space_ray                      core/rendering/integrator/volume/tracking.inl:12-29
//   phase_hg, sample_phase                 core/scene/material/material.inl:15-40
//   similarity_relation_scale              core/scene/material/material.cpp:108-135 (range 16-64)
//   AABB::intersect_inside                 base/math/aabb.inl:184-226
//   offset_f                               core/scene/ray_offset.inl:31-34
#pragma once

#include "bsdf.hpp"
#include "common.hpp"
#include "device_math.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace mi {

// CM, core/scene/material/collision_coefficients.hpp:17-35
struct Medium_cm {
    float minorant_mu_a;
    float minorant_mu_s;
    float majorant_mu_a;
    float majorant_mu_s;
};

struct HostGridtree {
    std::vector<uint32_t>  nodes;  // has_children : 1 | children_or_data : 31 (bit 0 = has_children)
    std::vector<Medium_cm> data;
    int32_t                num_cells[3] = {0, 0, 0};
};

// Octree_builder::build for a one-channel density grid (x fastest) and the
// medium's coefficients per unit density (rgb).  uniform: every top cell
// split to its 8^3-voxel leaves (a flat majorant grid).
void build_gridtree(float const* density, int32_t const dims[3], float const cc_a[3],
                    float const cc_s[3], HostGridtree& out, bool uniform = false);

namespace dev {

inline constexpr int32_t  Log2_cell_dim = 5;
inline constexpr int32_t  Cell_dim      = 1 << Log2_cell_dim;
inline constexpr uint32_t Grid_empty    = 0x7FFFFFFFu;
inline constexpr float    Min_mt        = 1e-10f;
inline constexpr float    Abort_epsilon = 7.5e-4f;
inline constexpr uint32_t SR_low        = 16;
inline constexpr uint32_t SR_high       = 64;

struct Medium_view {
    float const*     density;
    int32_t          dims[3];
    float            inv_dims[3];
    uint32_t         num_cells[3];
    uint32_t const*  nodes;
    Medium_cm const* data;
    float            box_lo[3];
    float            iext[3];  // 1 / extent of the box
    float            cc_a[3];
    float            cc_s[3];
    float            g;
    uint32_t         present;
};

__device__ __forceinline__ float saturate(float x) {
    return fminf(fmaxf(x, 0.f), 1.f);
}

// Nearest_3D<clamp>::sample_1 (texture_sampler.cpp:124-172)
__device__ inline float density_nearest(Medium_view const& m, V3 uvw) {
    int32_t const x = min(int32_t(saturate(uvw.x) * float(m.dims[0])), m.dims[0] - 1);
    int32_t const y = min(int32_t(saturate(uvw.y) * float(m.dims[1])), m.dims[1] - 1);
    int32_t const z = min(int32_t(saturate(uvw.z) * float(m.dims[2])), m.dims[2] - 1);
    return m.density[(size_t(z) * size_t(m.dims[1]) + size_t(y)) * size_t(m.dims[0]) + size_t(x)];
}

// Linear_3D<clamp>::sample_1 (texture_sampler.cpp:175-188, 256-282)
__device__ inline float density(Medium_view const& m, V3 uvw) {
    float const u  = saturate(uvw.x) * float(m.dims[0]) - 0.5f;
    float const v  = saturate(uvw.y) * float(m.dims[1]) - 0.5f;
    float const w  = saturate(uvw.z) * float(m.dims[2]) - 0.5f;
    float const fu = floorf(u);
    float const fv = floorf(v);
    float const fw = floorf(w);
    int32_t const x = int32_t(fu);
    int32_t const y = int32_t(fv);
    int32_t const z = int32_t(fw);
    int32_t const bx = m.dims[0] - 1;
    int32_t const by = m.dims[1] - 1;
    int32_t const bz = m.dims[2] - 1;
    int32_t const x0 = x < 0 ? 0 : x;
    int32_t const y0 = y < 0 ? 0 : y;
    int32_t const z0 = z < 0 ? 0 : z;
    int32_t const x1 = x >= bx ? bx : x + 1;
    int32_t const y1 = y >= by ? by : y + 1;
    int32_t const z1 = z >= bz ? bz : z + 1;
    float const s  = u - fu;
    float const t  = v - fv;
    float const q  = w - fw;
    size_t const sx = size_t(m.dims[0]);
    size_t const sxy = sx * size_t(m.dims[1]);
    auto at = [&](int32_t a, int32_t b, int32_t c) { return m.density[size_t(c) * sxy + size_t(b) * sx + size_t(a)]; };
    float const _s = 1.f - s;
    float const _t = 1.f - t;
    float const c0 = _t * (_s * at(x0, y0, z0) + s * at(x1, y0, z0)) + t * (_s * at(x0, y1, z0) + s * at(x1, y1, z0));
    float const c1 = _t * (_s * at(x0, y0, z1) + s * at(x1, y0, z1)) + t * (_s * at(x0, y1, z1) + s * at(x1, y1, z1));
    return lerpf(c0, c1, q);
}

__device__ __forceinline__ float offset_f(float t) {
    return t < (1.f / 32.f) ? t + (1.f / 65536.f) : __int_as_float(__float_as_int(t) + 256);
}

// A ray in the texture space of the medium box: t keeps world distance.
struct Tray {
    V3    o;
    V3    d;
    V3    inv;
    float min_t;
    float max_t;

    __device__ V3 point(float t) const { return v3(o.x + t * d.x, o.y + t * d.y, o.z + t * d.z); }
};

// AABB::intersect_inside (box given in texture space)
__device__ inline bool intersect_inside(V3 lo, V3 hi, Tray const& r, float& hit_t) {
    float const l1x = (lo.x - r.o.x) * r.inv.x, l2x = (hi.x - r.o.x) * r.inv.x;
    float const l1y = (lo.y - r.o.y) * r.inv.y, l2y = (hi.y - r.o.y) * r.inv.y;
    float const l1z = (lo.z - r.o.z) * r.inv.z, l2z = (hi.z - r.o.z) * r.inv.z;
    float const inf_ = __int_as_float(0x7f800000);
    // minps / maxps semantics: a < b ? a : b, NaN lanes filtered as the reference does
    auto mn = [](float a, float b) { return a < b ? a : b; };
    auto mx = [](float a, float b) { return a > b ? a : b; };
    float const fa1x = mn(l1x, inf_), fa2x = mn(l2x, inf_), fb1x = mx(l1x, -inf_), fb2x = mx(l2x, -inf_);
    float const fa1y = mn(l1y, inf_), fa2y = mn(l2y, inf_), fb1y = mx(l1y, -inf_), fb2y = mx(l2y, -inf_);
    float const fa1z = mn(l1z, inf_), fa2z = mn(l2z, inf_), fb1z = mx(l1z, -inf_), fb2z = mx(l2z, -inf_);
    float const max_x = mx(fa1x, fa2x), max_y = mx(fa1y, fa2y), max_z = mx(fa1z, fa2z);
    float const min_x = mn(fb1x, fb2x), min_y = mn(fb1y, fb2y), min_z = mn(fb1z, fb2z);
    float const max_out = mn(mn(max_x, max_y), max_z);
    float const min_out = mx(mx(min_x, min_y), min_z);
    hit_t = min_out < r.min_t ? max_out : min_out;
    return max_out >= r.min_t && r.min_t >= min_out && r.max_t >= min_out && max_out >= min_out;
}

// Gridtree::intersect: clips r.max_t to the leaf cell around r.point(min_t)
__device__ inline bool gridtree_intersect(Medium_view const& m, Tray& r, Medium_cm& cm) {
    V3 const      p  = r.point(r.min_t);
    int32_t const cx = int32_t(float(m.dims[0]) * p.x);
    int32_t const cy = int32_t(float(m.dims[1]) * p.y);
    int32_t const cz = int32_t(float(m.dims[2]) * p.z);
    int32_t const vx = cx >> Log2_cell_dim;
    int32_t const vy = cy >> Log2_cell_dim;
    int32_t const vz = cz >> Log2_cell_dim;
    if (uint32_t(vx) >= m.num_cells[0] || uint32_t(vy) >= m.num_cells[1] ||
        uint32_t(vz) >= m.num_cells[2]) {
        return false;
    }
    uint32_t index = (uint32_t(vz) * m.num_cells[1] + uint32_t(vy)) * m.num_cells[0] + uint32_t(vx);
    int32_t  b0[3] = {vx << Log2_cell_dim, vy << Log2_cell_dim, vz << Log2_cell_dim};
    int32_t  b1[3] = {b0[0] + Cell_dim, b0[1] + Cell_dim, b0[2] + Cell_dim};
    int32_t const c[3] = {cx, cy, cz};
    for (;;) {
        uint32_t const node = m.nodes[index];
        if (0 == (node & 1u)) {
            break;
        }
        index = node >> 1;
        for (int k = 0; k < 3; ++k) {
            int32_t const half   = (b1[k] - b0[k]) >> 1;
            int32_t const center = b0[k] + half;
            if (c[k] < center) {
                b1[k] = center;
            } else {
                b0[k] = center;
                index += 1u << k;
            }
        }
    }
    V3 const lo = v3(float(b0[0]) * m.inv_dims[0], float(b0[1]) * m.inv_dims[1], float(b0[2]) * m.inv_dims[2]);
    V3 const hi = v3(float(b1[0]) * m.inv_dims[0], float(b1[1]) * m.inv_dims[1], float(b1[2]) * m.inv_dims[2]);
    float hit_t;
    if (intersect_inside(lo, hi, r, hit_t)) {
        if (r.max_t > hit_t) {
            r.max_t = hit_t;
        }
    } else {
        r.max_t = r.min_t;
        return false;
    }
    uint32_t const node = m.nodes[index];
    if (Grid_empty == (node >> 1)) {
        return false;
    }
    cm = m.data[node >> 1];
    return true;
}

// van_de_hulst(g, gs) with gs lerped towards 0 over the range
__device__ __forceinline__ float similarity_scale(float g, uint32_t depth) {
    float gs;
    if (depth < SR_low) {
        gs = g;
    } else if (depth < SR_high) {
        gs = lerpf(g, 0.f, (1.f / float(SR_high - SR_low)) * float(depth - SR_low));
    } else {
        gs = 0.f;
    }
    return (1.f - g) / (1.f - gs);
}

__device__ __forceinline__ float phase_hg(float cos_theta, float g) {
    float const gg    = g * g;
    float const denom = 1.f + gg + 2.f * g * cos_theta;
    return (1.f / (4.f * Pi)) * (1.f - gg) / (denom * sqrtf(denom));
}

// sample_phase: direction and pdf (= phase value)
__device__ inline V3 sample_phase(V3 wo, float r0, float r1, float g, float& pdf) {
    float cos_theta;
    if (fabsf(g) < 0.001f) {
        cos_theta = 1.f - 2.f * r0;
    } else {
        float const gg  = g * g;
        float const sqr = (1.f - gg) / (1.f - g + 2.f * g * r0);
        cos_theta       = (1.f + gg - sqr * sqr) / (2.f * g);
    }
    float const sin_theta = sqrtf(fmaxf(0.f, 1.f - cos_theta * cos_theta));
    float const phi       = r1 * (2.f * Pi);
    Layer const l         = frame_from_normal(wo);
    float       sp, cp;
    sincosf(phi, &sp, &cp);
    V3 const z = -wo;
    V3 const wi = (sin_theta * cp * l.t + sin_theta * sp * l.b) + cos_theta * z;
    pdf         = phase_hg(-cos_theta, g);
    return wi;
}

}  // namespace dev

}  // namespace mi

namespace mi::dev {

// Per-lane random numbers for the tracking loops (the path's generator).
template <typename Rng>
__device__ __forceinline__ float medium_uniform(Rng& rng) {
    return rng.uniform();
}

// Collision coefficients per unit density of the channels a path carries:
// rgb (the medium's own) or, for hero-wavelength paths, the four wavelengths'
// upsampled values.
template <int NC>
struct Channel_cc {
    float a[NC];
    float s[NC];
};

__device__ __forceinline__ Channel_cc<3> rgb_channels(Medium_view const& m) {
    Channel_cc<3> cc;
    for (int c = 0; c < 3; ++c) {
        cc.a[c] = m.cc_a[c];
        cc.s[c] = m.cc_s[c];
    }
    return cc;
}

template <int NC>
__device__ __forceinline__ float average_channels(float const v[NC]) {
    if constexpr (3 == NC) {
        return (v[0] + v[1] + v[2]) / 3.f;
    } else {
        return (((v[0] + v[1]) + v[2]) + v[3]) * 0.25f;
    }
}

// collision coefficients at uvw, mu_s scaled by srs; nearest: the filter
// the path tracer asks for after its second vertex (Filter::Nearest,
// pathtracer_mis.cpp:121), else the grid's linear sampler
template <int NC>
__device__ __forceinline__ void coefficients(Medium_view const& m, Channel_cc<NC> const& cc, V3 uvw,
                                             float srs, bool nearest, float a[NC], float s[NC]) {
    float const dn = nearest ? density_nearest(m, uvw) : density(m, uvw);
    for (int c = 0; c < NC; ++c) {
        a[c] = dn * cc.a[c];
        s[c] = (dn * cc.s[c]) * srs;
    }
}

// Tracking::tracking (scattering, no emission): true = scattered at t_out.
// The channel weights are history-aware (the scatter / null probabilities
// average the channels weighted by the path so far).
template <int NC, typename Rng>
__device__ inline bool tracking(Medium_view const& m, Channel_cc<NC> const& cc, Tray const& r,
                                Medium_cm const& cm, float srs, bool nearest, Rng& rng,
                                float& t_out, float w[NC]) {
    float const mt = cm.majorant_mu_a + cm.majorant_mu_s;
    if (mt < Min_mt) {
        return false;
    }
    float lw[NC];
    for (int c = 0; c < NC; ++c) {
        lw[c] = w[c];
    }
    float const imt = 1.f / mt;
    for (float t = r.min_t, d = r.max_t;;) {
        float const r0 = medium_uniform(rng);
        t -= logf(1.f - r0) * imt;
        if (t > d) {
            for (int c = 0; c < NC; ++c) {
                w[c] = lw[c];
            }
            return false;
        }
        float mu_a[NC], mu_s[NC];
        coefficients<NC>(m, cc, r.point(t), srs, nearest, mu_a, mu_s);
        float mu_n[NC], msw[NC], mnw[NC];
        for (int c = 0; c < NC; ++c) {
            mu_n[c] = mt - (mu_a[c] + mu_s[c]);
            msw[c]  = mu_s[c] * lw[c];
            mnw[c]  = mu_n[c] * lw[c];
        }
        float const ms = average_channels<NC>(msw);
        float const mn = average_channels<NC>(mnw);
        float const ic = 1.f / (ms + mn);
        float const ps = ms * ic;
        float const pn = mn * ic;
        float const r1 = medium_uniform(rng);
        if ((r1 <= 1.f - pn) & (ps > 0.f)) {
            for (int c = 0; c < NC; ++c) {
                w[c] = lw[c] * (mu_s[c] / (mt * ps));
            }
            t_out = t;
            return true;
        }
        for (int c = 0; c < NC; ++c) {
            lw[c] *= mu_n[c] / (mt * pn);
        }
    }
}

template <int NC>
__device__ __forceinline__ bool all_below(float const w[NC], float e) {
    bool below = true;
    for (int c = 0; c < NC; ++c) {
        below = below && w[c] < e;
    }
    return below;
}

// tracking_transmitted with its residual-ratio-tracking branch
template <int NC, typename Rng>
__device__ inline bool tracking_transmitted(Medium_view const& m, Channel_cc<NC> const& cc,
                                            Tray const& r, Medium_cm const& cm, float srs,
                                            bool nearest, Rng& rng, float w[NC]) {
    float const mt = cm.majorant_mu_a + cm.majorant_mu_s;
    if (mt < Min_mt) {
        return true;
    }
    float const minorant = cm.minorant_mu_a + cm.minorant_mu_s;
    float       control  = 0.f;
    float       mr       = mt;
    if (minorant > 0.f) {
        float const att = expf(-(r.max_t - r.min_t) * minorant);
        for (int c = 0; c < NC; ++c) {
            w[c] *= att;
        }
        if (all_below<NC>(w, Abort_epsilon)) {
            return false;
        }
        mr = mt - minorant;
        if (mr < Min_mt) {
            return true;
        }
        control = minorant;
    }
    float const imt = 1.f / mr;
    for (float t = r.min_t, d = r.max_t;;) {
        float const r0 = medium_uniform(rng);
        t -= logf(1.f - r0) * imt;
        if (t > d) {
            return true;
        }
        float mu_a[NC], mu_s[NC];
        coefficients<NC>(m, cc, r.point(t), srs, nearest, mu_a, mu_s);
        for (int c = 0; c < NC; ++c) {
            float const mu_t = (mu_a[c] + mu_s[c]) - control;
            w[c] *= imt * (mr - mu_t);
        }
        if (all_below<NC>(w, Abort_epsilon)) {
            return false;
        }
    }
}

// Texture-space ray of a world segment [t0, t1] (the part inside the box).
__device__ inline Tray texture_ray(Medium_view const& m, V3 o, V3 d, float t0, float t1) {
    Tray r;
    r.o     = v3((o.x - m.box_lo[0]) * m.iext[0], (o.y - m.box_lo[1]) * m.iext[1],
                 (o.z - m.box_lo[2]) * m.iext[2]);
    r.d     = v3(d.x * m.iext[0], d.y * m.iext[1], d.z * m.iext[2]);
    r.inv   = v3(reciprocal_lane(r.d.x), reciprocal_lane(r.d.y), reciprocal_lane(r.d.z));
    r.min_t = t0;
    r.max_t = t1;
    return r;
}

// World-space overlap [t0, t1] of a ray with the medium box within
// [min_t, max_t].  A ray that starts outside the box is tracked from
// offset_f of its entry distance, as a ray continued through a medium's
// boundary surface is (pathtracer_mis.cpp:185-186).
__device__ inline bool box_overlap(Medium_view const& m, V3 o, V3 d, float min_t, float max_t,
                                   float& t0, float& t1) {
    float const lo[3] = {m.box_lo[0], m.box_lo[1], m.box_lo[2]};
    float const hi[3] = {m.box_lo[0] + 1.f / m.iext[0], m.box_lo[1] + 1.f / m.iext[1],
                         m.box_lo[2] + 1.f / m.iext[2]};
    float const oo[3] = {o.x, o.y, o.z};
    float const dd[3] = {d.x, d.y, d.z};
    float       a     = min_t;
    float       b     = max_t;
    for (int k = 0; k < 3; ++k) {
        if (0.f == dd[k]) {
            if (oo[k] < lo[k] || oo[k] > hi[k]) {
                return false;
            }
            continue;
        }
        float const inv = reciprocal_lane(dd[k]);
        float       n   = (lo[k] - oo[k]) * inv;
        float       f   = (hi[k] - oo[k]) * inv;
        if (n > f) {
            float const x = n;
            n             = f;
            f             = x;
        }
        // NaN (0 * inf) leaves the interval as it is
        a = n > a ? n : a;
        b = f < b ? f : b;
    }
    t0 = a > min_t ? offset_f(a) : a;
    t1 = b;
    return t0 < t1;
}

// Tracking_multi::integrate for one segment (tracking_multi.cpp:101-162):
// 1 scattered at t_out, 0 passed, 2 aborted (every channel of w below
// Abort_epsilon, scattered or not).  w multiplies the throughput.
template <int NC, typename Rng>
__device__ inline int track_segment(Medium_view const& m, Channel_cc<NC> const& cc, V3 o, V3 d,
                                    float t0, float t1, uint32_t depth, bool nearest, Rng& rng,
                                    float& t_out, float w[NC]) {
    Tray        r     = texture_ray(m, o, d, t0, t1);
    float const srs   = similarity_scale(m.g, depth);
    int         event = 0;
    for (int c = 0; c < NC; ++c) {
        w[c] = 1.f;
    }
    for (float const end = t1; r.min_t < end;) {
        Medium_cm cm;
        if (gridtree_intersect(m, r, cm)) {
            cm.minorant_mu_s *= srs;
            cm.majorant_mu_s *= srs;
            if (tracking<NC>(m, cc, r, cm, srs, nearest, rng, t_out, w)) {
                event = 1;
                break;
            }
        }
        r.min_t = offset_f(r.max_t);
        r.max_t = end;
    }
    return all_below<NC>(w, Abort_epsilon) ? 2 : event;
}

template <typename Rng>
__device__ inline int track_segment(Medium_view const& m, V3 o, V3 d, float t0, float t1,
                                    uint32_t depth, bool nearest, Rng& rng, float& t_out,
                                    float w[3]) {
    return track_segment<3>(m, rgb_channels(m), o, d, t0, t1, depth, nearest, rng, t_out, w);
}

// Tracking::transmittance for the part of a segment inside the box:
// false = blocked (tr below Abort_epsilon), else tr.
template <int NC, typename Rng>
__device__ inline bool transmittance(Medium_view const& m, Channel_cc<NC> const& cc, V3 o, V3 d,
                                     float t0, float t1, uint32_t depth, bool nearest, Rng& rng,
                                     float tr[NC]) {
    for (int c = 0; c < NC; ++c) {
        tr[c] = 1.f;
    }
    if (offset_f(t0) >= t1) {
        return true;
    }
    Tray        r   = texture_ray(m, o, d, t0, t1);
    float const srs = similarity_scale(m.g, depth);
    for (float const end = t1; r.min_t < end;) {
        Medium_cm cm;
        if (gridtree_intersect(m, r, cm)) {
            cm.minorant_mu_s *= srs;
            cm.majorant_mu_s *= srs;
            if (!tracking_transmitted<NC>(m, cc, r, cm, srs, nearest, rng, tr)) {
                for (int c = 0; c < NC; ++c) {
                    tr[c] = 0.f;
                }
                return false;
            }
        }
        r.min_t = offset_f(r.max_t);
        r.max_t = end;
    }
    return true;
}

template <typename Rng>
__device__ inline bool transmittance(Medium_view const& m, V3 o, V3 d, float t0, float t1,
                                     uint32_t depth, bool nearest, Rng& rng, float tr[3]) {
    return transmittance<3>(m, rgb_channels(m), o, d, t0, t1, depth, nearest, rng, tr);
}

}  // namespace mi::dev
