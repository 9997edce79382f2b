// The arithmetic of sprout's light-tree builder (Tree_builder,
// core/scene/light/light_tree_builder.cpp:18-760), shared by the host
// builder (light_tree.cpp) and the device builder (light_build.hip) so that
// both produce the same tree bit for bit:
//
//   cone_cost                  light_tree_builder.cpp:74-83
//   variance                   light_tree_builder.cpp:85-101
//   Split_candidate::evaluate  light_tree_builder.cpp:149-234 (scene variant)
//   cone::merge                base/math/cone.inl:17-60
//   set_rotation, transform    base/math/matrix3x3.inl:49-53, 106-132
//
// min / max keep std::min / std::max's comparisons.  The trigonometry is
// evaluated in double and rounded to float: the host's libm and the device's
// give the same float there (their double results lie within an ulp of the
// exact value, so they round alike except when that value sits within an
// ulp of a float rounding boundary), where their float functions differ in
// one result out of ten.  Against sprout's float std::acos / cos / sin the
// values move by at most an ulp.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace mi::lb {

inline constexpr float    Pi                    = 3.14159265358979323846f;
inline constexpr uint32_t Scene_sweep_threshold = 128;
inline constexpr uint32_t Part_sweep_threshold  = 32;
inline constexpr uint32_t Num_slices            = 16;

__host__ __device__ inline float acos_r(float x) {
    return float(::acos(double(x)));
}
__host__ __device__ inline float cos_r(float x) {
    return float(::cos(double(x)));
}
__host__ __device__ inline float sin_r(float x) {
    return float(::sin(double(x)));
}

// std::min / std::max
__host__ __device__ inline float smin(float a, float b) {
    return (b < a) ? b : a;
}
__host__ __device__ inline float smax(float a, float b) {
    return (a < b) ? b : a;
}

__host__ __device__ inline float dot3(float const* a, float const* b) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

struct Box {
    float lo[3] = {3.40282347e+38f, 3.40282347e+38f, 3.40282347e+38f};
    float hi[3] = {-3.40282347e+38f, -3.40282347e+38f, -3.40282347e+38f};

    __host__ __device__ void merge(float const* blo, float const* bhi) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = smin(lo[k], blo[k]);
            hi[k] = smax(hi[k], bhi[k]);
        }
    }
    __host__ __device__ void merge(Box const& b) { merge(b.lo, b.hi); }
    __host__ __device__ float extent(int k) const { return hi[k] - lo[k]; }
    __host__ __device__ float position(int k) const { return 0.5f * (lo[k] + hi[k]); }
    __host__ __device__ float surface_area() const {
        float const d0 = extent(0), d1 = extent(1), d2 = extent(2);
        return 2.f * (d0 * d1 + d0 * d2 + d1 * d2);
    }
};

struct Cone {
    float v[4] = {1.f, 1.f, 1.f, 1.f};  // float4(1) is the empty cone
};

__host__ __device__ inline bool same_bits(Cone const& a, Cone const& b) {
    for (int k = 0; k < 4; ++k) {
        uint32_t x, y;
        __builtin_memcpy(&x, &a.v[k], 4);
        __builtin_memcpy(&y, &b.v[k], 4);
        if (x != y) {
            return false;
        }
    }
    return true;
}

// cone::merge, base/math/cone.inl:17-60
__host__ __device__ inline Cone merge(Cone a, Cone b) {
    if (1.f == a.v[0] && 1.f == a.v[1] && 1.f == a.v[2] && 1.f == a.v[3]) {
        return b;
    }
    if (same_bits(a, b)) {
        return a;
    }
    float a_angle = acos_r(a.v[3]);
    float b_angle = acos_r(b.v[3]);
    if (b_angle > a_angle) {
        Cone const t = a;
        a            = b;
        b            = t;
        float const s = a_angle;
        a_angle       = b_angle;
        b_angle       = s;
    }
    float const d_angle = acos_r(smin(smax(dot3(a.v, b.v), -1.f), 1.f));
    if (smin(d_angle + b_angle, Pi) <= a_angle) {
        return a;
    }
    float const o_angle = (a_angle + d_angle + b_angle) / 2.f;
    // (the reference computes, and drops, a full-sphere cone when o_angle >= pi)
    float const r_angle = o_angle - a_angle;

    float c[3] = {a.v[1] * b.v[2] - a.v[2] * b.v[1], a.v[2] * b.v[0] - a.v[0] * b.v[2],
                  a.v[0] * b.v[1] - a.v[1] * b.v[0]};
    float const cc = dot3(c, c);
    if (!(cc > 0.f)) {
        // opposite axes: the reference normalizes a zero vector here and gets
        // a NaN cone; the cone that bounds both is the whole sphere
        Cone r = a;
        r.v[3] = -1.f;
        return r;
    }
    float const cl = 1.f / sqrtf(cc);
    for (float& x : c) {
        x *= cl;
    }
    // set_rotation(rot, c, r_angle)
    float const cs = cos_r(r_angle);
    float const sn = sin_r(r_angle);
    float const t  = 1.f - cs;
    float       m[3][3];
    m[0][0]    = cs + c[0] * c[0] * t;
    m[1][1]    = cs + c[1] * c[1] * t;
    m[2][2]    = cs + c[2] * c[2] * t;
    float tmp1 = c[0] * c[1] * t;
    float tmp2 = c[2] * sn;
    m[1][0]    = tmp1 + tmp2;
    m[0][1]    = tmp1 - tmp2;
    tmp1       = c[0] * c[2] * t;
    tmp2       = c[1] * sn;
    m[2][0]    = tmp1 - tmp2;
    m[0][2]    = tmp1 + tmp2;
    tmp1       = c[1] * c[2] * t;
    tmp2       = c[0] * sn;
    m[2][1]    = tmp1 + tmp2;
    m[1][2]    = tmp1 - tmp2;
    // transform_vector(rot, a.xyz): row vector times matrix
    float ax[3];
    for (int j = 0; j < 3; ++j) {
        ax[j] = a.v[0] * m[0][j] + a.v[1] * m[1][j] + a.v[2] * m[2][j];
    }
    float const al = 1.f / sqrtf(dot3(ax, ax));
    Cone        r;
    r.v[0] = ax[0] * al;
    r.v[1] = ax[1] * al;
    r.v[2] = ax[2] * al;
    r.v[3] = cos_r(o_angle);
    return r;
}

__host__ __device__ inline float cone_cost(float cos) {
    float const o   = acos_r(cos);
    float const w   = smin(o + (Pi / 2.f), Pi);
    float const sin = sin_r(o);
    float const b   = (Pi / 2.f) * (2.f * w * sin - cos_r(o - 2.f * w) - 2.f * o * sin + cos);
    return (2.f * Pi) * (1.f - cos) + b;
}

// The lights as the builder reads them: box, cone, power, two-sidedness.
struct Light_set {
    float const*   lo;         // 3 per light
    float const*   hi;         // 3 per light
    float const*   cones;      // 4 per light
    float const*   powers;
    uint8_t const* two_sided;  // or null (one-sided)

    __host__ __device__ Cone cone(uint32_t l) const {
        Cone c;
        for (int k = 0; k < 4; ++k) {
            c.v[k] = cones[4 * size_t(l) + k];
        }
        return c;
    }
    __host__ __device__ bool two(uint32_t l) const { return two_sided && 0 != two_sided[l]; }
};

// variance over mapping[begin, end) of the lights with power
__host__ __device__ inline float variance(uint32_t const* mapping, uint32_t begin, uint32_t end,
                                          Light_set const& set) {
    float ap  = 0.f;
    float aps = 0.f;
    for (uint32_t i = begin, n = 0; i < end; ++i) {
        float const p = set.powers[mapping[i]];
        if (p > 0.f) {
            float const in = 1.f / float(++n);
            ap += (p - ap) * in;
            aps += (p * p - aps) * in;
        }
    }
    return fabsf(aps - ap * ap);
}

// A split plane and what it makes of a node's lights.
struct Candidate {
    Box   box0, box1;
    Cone  cone0, cone1;
    float power0 = 0.f, power1 = 0.f;
    float d      = 0.f;
    float cost   = 0.f;
    int   axis   = 0;
    bool  two0 = false, two1 = false;
    bool  exhausted = false;

    __host__ __device__ bool behind(float const* hi) const { return hi[axis] < d; }

    // Split_candidate::evaluate over mapping[begin, end) of a node with
    // the given bounds and cone weight
    __host__ __device__ void evaluate(uint32_t begin, uint32_t end, uint32_t const* mapping,
                                      Box const& bounds, float cone_weight, Light_set const& set) {
        uint32_t n0 = 0, n1 = 0;
        Box      b0, b1;
        Cone     c0, c1;
        bool     t0 = false, t1 = false;
        float    p0 = 0.f, p1 = 0.f;
        for (uint32_t i = begin; i < end; ++i) {
            uint32_t const l     = mapping[i];
            float const    power = set.powers[l];
            if (0.f == power) {
                continue;
            }
            float const* lo = set.lo + 3 * size_t(l);
            float const* hi = set.hi + 3 * size_t(l);
            if (behind(hi)) {
                ++n0;
                b0.merge(lo, hi);
                c0 = merge(c0, set.cone(l));
                t0 = t0 || set.two(l);
                p0 += power;
            } else {
                ++n1;
                b1.merge(lo, hi);
                c1 = merge(c1, set.cone(l));
                t1 = t1 || set.two(l);
                p1 += power;
            }
        }
        float const e0  = bounds.extent(0), e1 = bounds.extent(1), e2 = bounds.extent(2);
        float const reg = smax(smax(e0, e1), e2) / bounds.extent(axis);
        float const sa  = bounds.surface_area();
        box0 = b0, box1 = b1, cone0 = c0, cone1 = c1, power0 = p0, power1 = p1, two0 = t0, two1 = t1;
        if (0 == n0 || 0 == n1) {
            cost      = 2.f * reg * (p0 + p1) * (4.f * Pi) * sa * float(end - begin);
            exhausted = true;
        } else {
            float const wa = cone_cost(c0.v[3]) * (t0 ? 2.f : 1.f);
            float const wb = cone_cost(c1.v[3]) * (t1 ? 2.f : 1.f);
            cost = reg * (((p0 * wa * b0.surface_area()) + (p1 * wb * b1.surface_area())) /
                          (sa * cone_weight));
            exhausted = false;
        }
    }
};

// The number of split candidates of a node of len lights (sweep below the
// threshold: three per light; else the 16-slice grid of
// evaluate_splits, light_tree_builder.cpp:346-420).
// A node whose lights all sit at one point has no grid step: no candidates.
__host__ __device__ inline uint32_t grid_steps(Box const& bounds, int a) {
    float const e0 = bounds.extent(0), e1 = bounds.extent(1), e2 = bounds.extent(2);
    int const   la = e0 >= e1 ? (e0 >= e2 ? 0 : 2) : (e1 >= e2 ? 1 : 2);
    float const step = bounds.extent(la) / float(Num_slices);
    if (!(step > 0.f)) {
        return 0;
    }
    return uint32_t(ceilf(bounds.extent(a) / step));
}

__host__ __device__ inline uint32_t grid_candidates(Box const& bounds) {
    uint32_t n = 0;
    for (int a = 0; a < 3; ++a) {
        uint32_t const s = grid_steps(bounds, a);
        n += s > 0 ? s - 1 : 0;
    }
    return n;
}

// Candidate k of the grid (k counts the slices of axis 0, then 1, then 2).
__host__ __device__ inline void grid_candidate(Box const& bounds, uint32_t k, Candidate& c) {
    for (int a = 0; a < 3; ++a) {
        uint32_t const num_steps = grid_steps(bounds, a);
        uint32_t const n         = num_steps > 0 ? num_steps - 1 : 0;
        if (k < n) {
            float const step_a = bounds.extent(a) / float(num_steps);
            c.d                = bounds.lo[a] + float(k + 1) * step_a;
            c.axis             = a;
            return;
        }
        k -= n;
    }
}

// Candidate k of a node's sweep: light mapping[begin + k / 3]'s box max on axis k % 3.
__host__ __device__ inline void sweep_candidate(uint32_t const* mapping, uint32_t begin,
                                                uint32_t k, Light_set const& set, Candidate& c) {
    uint32_t const l = mapping[begin + k / 3];
    c.axis           = int(k % 3);
    c.d              = set.hi[3 * size_t(l) + c.axis];
}

// One emissive triangle as a light (a, b, c its vertices): its box, centre,
// half diagonal, unit normal and area.  False for a triangle of no area.
struct Tri_light {
    float lo[3], hi[3], center[3], n[3];
    float radius, area;
};

__host__ __device__ inline bool triangle_light(float const* a, float const* b, float const* c,
                                               Tri_light& t) {
    float const e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    float const e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    float       n[3]  = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                         e1[0] * e2[1] - e1[1] * e2[0]};
    float const len   = sqrtf((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
    t.area            = 0.5f * len;
    if (!(t.area > 0.f)) {
        return false;
    }
    for (int k = 0; k < 3; ++k) {
        t.n[k]      = n[k] / len;
        t.lo[k]     = smin(smin(a[k], b[k]), c[k]);
        t.hi[k]     = smax(smax(a[k], b[k]), c[k]);
        t.center[k] = 0.5f * (t.lo[k] + t.hi[k]);
    }
    float const ext[3] = {t.hi[0] - t.lo[0], t.hi[1] - t.lo[1], t.hi[2] - t.lo[2]};
    t.radius           = 0.5f * sqrtf((ext[0] * ext[0] + ext[1] * ext[1]) + ext[2] * ext[2]);
    return true;
}

// Part::init (triangle_mesh.cpp:177-239) over a part's triangles in order:
// the box, the power-weighted dominant axis, the widest angle from it
// (its cosine closes the cone) and the total power (area).
struct Part_light {
    float lo[3], hi[3];
    float cone[4];
    float total;
};

template <class Tri>  // tri(i) -> Tri_light of the part's i-th triangle
__host__ __device__ inline void part_light(uint32_t n, Tri tri, Part_light& p) {
    float axis[3] = {0.f, 0.f, 0.f};
    p.total       = 0.f;
    for (int k = 0; k < 3; ++k) {
        p.lo[k] = __builtin_huge_valf();
        p.hi[k] = -__builtin_huge_valf();
    }
    for (uint32_t i = 0; i < n; ++i) {
        Tri_light const t = tri(i);
        for (int k = 0; k < 3; ++k) {
            p.lo[k] = smin(p.lo[k], t.lo[k]);
            p.hi[k] = smax(p.hi[k], t.hi[k]);
            axis[k] += t.area * t.n[k];
        }
        p.total += t.area;
    }
    float ax[3] = {axis[0] / p.total, axis[1] / p.total, axis[2] / p.total};
    float const al = sqrtf((ax[0] * ax[0] + ax[1] * ax[1]) + ax[2] * ax[2]);
    for (float& x : ax) {
        x /= al;
    }
    float angle = 0.f;
    for (uint32_t i = 0; i < n; ++i) {
        Tri_light const t  = tri(i);
        float const     cc = (ax[0] * t.n[0] + ax[1] * t.n[1]) + ax[2] * t.n[2];
        angle              = smax(angle, acos_r(smin(smax(cc, -1.f), 1.f)));
    }
    p.cone[0] = ax[0];
    p.cone[1] = ax[1];
    p.cone[2] = ax[2];
    p.cone[3] = cos_r(angle);
}

}  // namespace mi::lb
