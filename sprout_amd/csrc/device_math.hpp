// Device arithmetic shared by the traversal and build kernels.
//
// Every expression is written with the rounding the reference gets from
// g++ -O3 -march=native (x86-64 with FMA): products the compiler fuses are
// written with __builtin_fmaf, everything else rounds per operation.  The
// library is compiled with -ffp-contract=off so nothing else gets fused.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mi::dev {

inline constexpr uint32_t Null = 0xFFFFFFFFu;

__device__ __forceinline__ float inf() {
    return __builtin_inff();
}

// Pinned reciprocal of a scalar: base/math/scalar.inl:67-71 with the
// rcpss estimate replaced by the correctly rounded 1/x.  Scalar intrinsics,
// so the Newton step is not fused: (r + r) - (r * r) * x.
__device__ __forceinline__ float reciprocal(float x) {
    float const r  = 1.f / x;
    float const rr = r * r;
    float const m  = rr * x;
    return (r + r) - m;
}

// Pinned reciprocal3 lane: base/math/simd.inl:206-212 (vector intrinsics,
// which the compiler contracts to fnmadd): fma(-(r * r), x, r + r).
// Zero gives NaN (inf * 0), as in the reference.
__device__ __forceinline__ float reciprocal_lane(float x) {
    float const r  = 1.f / x;
    float const rr = r * r;
    return __builtin_fmaf(-rr, x, r + r);
}

struct F3 {
    float x, y, z;
};

// cross3 under AVX2: fmsub(b.yzx, a, a.yzx * b) permuted (base/math/simd.inl:173-188)
__device__ __forceinline__ F3 cross3(F3 a, F3 b) {
    return {__builtin_fmaf(a.y, b.z, -(a.z * b.y)), __builtin_fmaf(a.z, b.x, -(a.x * b.z)),
            __builtin_fmaf(a.x, b.y, -(a.y * b.x))};
}

// dot3_scalar: rounded products summed (x + y) + z (base/math/simd.inl:165-171)
__device__ __forceinline__ float dot3(F3 a, F3 b) {
    float const x = a.x * b.x;
    float const y = a.y * b.y;
    float const z = a.z * b.z;
    return (x + y) + z;
}

__device__ __forceinline__ F3 sub(F3 a, F3 b) {
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}

// Node::intersect_p (core/scene/bvh/scene_bvh_node.inl:133-163).  NaN lanes
// (inv_dir NaN times a zero distance) are filtered by min(l, +inf) and
// max(l, -inf) as there; fminf/fmaxf give the same values (up to the sign
// of zero, which no comparison below sees).
__device__ __forceinline__ bool slab(float4 lo, float4 hi, F3 o, F3 inv, float min_t, float max_t) {
    float const l1x = (lo.x - o.x) * inv.x;
    float const l1y = (lo.y - o.y) * inv.y;
    float const l1z = (lo.z - o.z) * inv.z;
    float const l2x = (hi.x - o.x) * inv.x;
    float const l2y = (hi.y - o.y) * inv.y;
    float const l2z = (hi.z - o.z) * inv.z;

    float const tmax_x = fmaxf(fminf(l1x, inf()), fminf(l2x, inf()));
    float const tmax_y = fmaxf(fminf(l1y, inf()), fminf(l2y, inf()));
    float const tmax_z = fmaxf(fminf(l1z, inf()), fminf(l2z, inf()));
    float const tmin_x = fminf(fmaxf(l1x, -inf()), fmaxf(l2x, -inf()));
    float const tmin_y = fminf(fmaxf(l1y, -inf()), fmaxf(l2y, -inf()));
    float const tmin_z = fminf(fmaxf(l1z, -inf()), fmaxf(l2z, -inf()));

    float const tmax = fminf(fminf(tmax_x, tmax_y), tmax_z);
    float const tmin = fmaxf(fmaxf(tmin_x, tmin_y), tmin_z);

    return (tmax >= min_t) & (max_t >= tmin) & (tmax >= tmin);
}

// slab with the box's entry distance tmin: while the ray's min_t and
// direction stay, the test at a smaller max_t' is exactly max_t' >= tmin
// (the other two terms do not involve max_t).  tmin is never NaN.
__device__ __forceinline__ bool slab_entry(float4 lo, float4 hi, F3 o, F3 inv, float min_t,
                                           float max_t, float& tmin_out) {
    float const l1x = (lo.x - o.x) * inv.x;
    float const l1y = (lo.y - o.y) * inv.y;
    float const l1z = (lo.z - o.z) * inv.z;
    float const l2x = (hi.x - o.x) * inv.x;
    float const l2y = (hi.y - o.y) * inv.y;
    float const l2z = (hi.z - o.z) * inv.z;

    float const tmax_x = fmaxf(fminf(l1x, inf()), fminf(l2x, inf()));
    float const tmax_y = fmaxf(fminf(l1y, inf()), fminf(l2y, inf()));
    float const tmax_z = fmaxf(fminf(l1z, inf()), fminf(l2z, inf()));
    float const tmin_x = fminf(fmaxf(l1x, -inf()), fmaxf(l2x, -inf()));
    float const tmin_y = fminf(fmaxf(l1y, -inf()), fmaxf(l2y, -inf()));
    float const tmin_z = fminf(fmaxf(l1z, -inf()), fmaxf(l2z, -inf()));

    float const tmax = fminf(fminf(tmax_x, tmax_y), tmax_z);
    float const tmin = fmaxf(fmaxf(tmin_x, tmin_y), tmin_z);
    tmin_out         = tmin;
    return (tmax >= min_t) & (max_t >= tmin) & (tmax >= tmin);
}

// triangle::intersect (core/scene/shape/triangle/triangle_primitive_mt.inl:61-93)
// with e1 = b - a and e2 = c - a precomputed (the same rounded subtraction).
// On a hit returns true and sets t, u, v.
__device__ __forceinline__ bool triangle(F3 a, F3 e1, F3 e2, F3 o, F3 d, float min_t, float max_t,
                                         float& t_out, float& u_out, float& v_out) {
    F3 const tv = sub(o, a);
    F3 const p  = cross3(d, e2);
    F3 const q  = cross3(tv, e1);

    float const e1_d_pv = dot3(e1, p);
    float const tv_d_pv = dot3(tv, p);
    float const di_d_qv = dot3(d, q);
    float const e2_d_qv = dot3(e2, q);

    float const inv_det = reciprocal(e1_d_pv);

    float const u  = tv_d_pv * inv_det;
    float const v  = di_d_qv * inv_det;
    float const t  = e2_d_qv * inv_det;
    float const uv = u + v;

    t_out = t;
    u_out = u;
    v_out = v;

    return (u >= 0.f) & (1.f >= u) & (v >= 0.f) & (1.f >= uv) & (t >= min_t) & (max_t >= t);
}

}  // namespace mi::dev
