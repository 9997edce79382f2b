// Host construction of the light tree: sprout's Tree_builder for the scene
// tree (core/scene/light/light_tree_builder.cpp:18-540), finite lights only.
//
//   cone_cost                  light_tree_builder.cpp:74-83
//   variance                   light_tree_builder.cpp:85-101
//   Split_candidate::evaluate  light_tree_builder.cpp:149-234 (scene variant)
//   evaluate_splits            light_tree_builder.cpp:346-420
//   split / serialize          light_tree_builder.cpp:583-640, 712-733
//   cone::merge                base/math/cone.inl:17-60
//   set_rotation, transform    base/math/matrix3x3.inl:49-53, 106-132
//
// Candidates are evaluated one after the other (the reference spreads large
// candidate sets over its thread pool; the result is the same).

#include "light_tree.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace mi {

namespace {

constexpr float    Pi                    = 3.14159265358979323846f;
constexpr uint32_t Scene_sweep_threshold = 128;
constexpr uint32_t Part_sweep_threshold  = 32;
constexpr uint32_t Num_slices            = 16;

struct F3 {
    float v[3];
    float&       operator[](int i) { return v[i]; }
    float const& operator[](int i) const { return v[i]; }
};

struct Box {
    F3 lo{{3.40282347e+38f, 3.40282347e+38f, 3.40282347e+38f}};
    F3 hi{{-3.40282347e+38f, -3.40282347e+38f, -3.40282347e+38f}};

    void merge(Box const& b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    F3 extent() const { return {{hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]}}; }
    F3 position() const {
        return {{0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2])}};
    }
    float surface_area() const {
        F3 const d = extent();
        return 2.f * (d[0] * d[1] + d[0] * d[2] + d[1] * d[2]);
    }
};

struct Cone {
    float v[4] = {1.f, 1.f, 1.f, 1.f};  // float4(1) is the empty cone
};

float dot3(float const* a, float const* b) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

// cone::merge, base/math/cone.inl:17-60
Cone merge(Cone a, Cone b) {
    if (1.f == a.v[0] && 1.f == a.v[1] && 1.f == a.v[2] && 1.f == a.v[3]) {
        return b;
    }
    if (0 == std::memcmp(a.v, b.v, sizeof(a.v))) {
        return a;
    }
    float a_angle = std::acos(a.v[3]);
    float b_angle = std::acos(b.v[3]);
    if (b_angle > a_angle) {
        std::swap(a, b);
        std::swap(a_angle, b_angle);
    }
    float const d_angle = std::acos(std::min(std::max(dot3(a.v, b.v), -1.f), 1.f));
    if (std::min(d_angle + b_angle, Pi) <= a_angle) {
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
    float const cl = 1.f / std::sqrt(cc);
    for (float& x : c) {
        x *= cl;
    }
    // set_rotation(rot, c, r_angle)
    float const cs = std::cos(r_angle);
    float const sn = std::sin(r_angle);
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
    float const al = 1.f / std::sqrt(dot3(ax, ax));
    Cone        r;
    r.v[0] = ax[0] * al;
    r.v[1] = ax[1] * al;
    r.v[2] = ax[2] * al;
    r.v[3] = std::cos(o_angle);
    return r;
}

float cone_cost(float cos) {
    float const o   = std::acos(cos);
    float const w   = std::min(o + (Pi / 2.f), Pi);
    float const sin = std::sin(o);
    float const b   = (Pi / 2.f) * (2.f * w * sin - std::cos(o - 2.f * w) - 2.f * o * sin + cos);
    return (2.f * Pi) * (1.f - cos) + b;
}

struct Lights {
    std::vector<Box>     boxes;
    std::vector<Cone>    cones;
    std::vector<float>   powers;
    std::vector<uint8_t> two_sided;
};

float variance(uint32_t const* lights, uint32_t begin, uint32_t end, Lights const& set) {
    float ap  = 0.f;
    float aps = 0.f;
    for (uint32_t i = begin, n = 0; i < end; ++i) {
        float const p = set.powers[lights[i]];
        if (p > 0.f) {
            float const in = 1.f / float(++n);
            ap += (p - ap) * in;
            aps += (p * p - aps) * in;
        }
    }
    return std::abs(aps - ap * ap);
}

struct Candidate {
    Box   box0, box1;
    Cone  cone0, cone1;
    float power0 = 0.f, power1 = 0.f;
    float d      = 0.f;
    float cost   = 0.f;
    int   axis   = 0;
    bool  two0 = false, two1 = false;
    bool  exhausted = false;

    bool behind(F3 const& p) const { return p[axis] < d; }

    void evaluate(uint32_t begin, uint32_t end, uint32_t const* lights, Box const& bounds,
                  float cone_weight, Lights const& set) {
        uint32_t n0 = 0, n1 = 0;
        Box      b0, b1;
        Cone     c0, c1;
        bool     t0 = false, t1 = false;
        float    p0 = 0.f, p1 = 0.f;
        for (uint32_t i = begin; i < end; ++i) {
            uint32_t const l     = lights[i];
            float const    power = set.powers[l];
            if (0.f == power) {
                continue;
            }
            if (behind(set.boxes[l].hi)) {
                ++n0;
                b0.merge(set.boxes[l]);
                c0 = merge(c0, set.cones[l]);
                t0 |= 0 != set.two_sided[l];
                p0 += power;
            } else {
                ++n1;
                b1.merge(set.boxes[l]);
                c1 = merge(c1, set.cones[l]);
                t1 |= 0 != set.two_sided[l];
                p1 += power;
            }
        }
        F3 const    extent = bounds.extent();
        float const reg    = std::max(std::max(extent[0], extent[1]), extent[2]) / extent[axis];
        float const sa     = bounds.surface_area();
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

struct Build_node {
    Box      bounds;
    Cone     cone;
    float    power      = 0.f;
    float    var        = 0.f;
    uint32_t middle     = 0;
    uint32_t children   = 0;
    uint32_t num_lights = 0;
    bool     two_sided  = false;
};

struct Builder {
    Lights const&           set;
    std::vector<uint32_t>&  mapping;
    std::vector<uint32_t>&  orders;
    std::vector<Build_node> nodes;
    std::vector<Candidate>  candidates;
    uint32_t                current = 1;
    uint32_t                order   = 0;
    // Primitive_tree of one part (light_tree_builder.cpp:539-758): leaves of
    // up to max_primitives lights, 32-light sweeps, exhausted nodes become
    // leaves, one-sided throughout
    bool     part           = false;
    uint32_t max_primitives = 4;

    Candidate evaluate_splits(uint32_t begin, uint32_t end, Box const& bounds, float cone_weight) {
        uint32_t const len = end - begin;
        candidates.clear();
        if (len < (part ? Part_sweep_threshold : Scene_sweep_threshold)) {
            for (uint32_t i = begin; i < end; ++i) {
                F3 const& mx = set.boxes[mapping[i]].hi;
                for (int a = 0; a < 3; ++a) {
                    Candidate c;
                    c.d    = mx[a];
                    c.axis = a;
                    candidates.push_back(c);
                }
            }
        } else {
            F3 const    position = bounds.position();
            F3 const    extent   = bounds.extent();
            F3 const    mn       = bounds.lo;
            int const   la       = extent[0] >= extent[1] ? (extent[0] >= extent[2] ? 0 : 2)
                                                          : (extent[1] >= extent[2] ? 1 : 2);
            float const step     = extent[la] / float(Num_slices);
            for (int a = 0; a < 3; ++a) {
                float const    extent_a  = extent[a];
                uint32_t const num_steps = uint32_t(std::ceil(extent_a / step));
                float const    step_a    = extent_a / float(num_steps);
                for (uint32_t i = 1; i < num_steps; ++i) {
                    F3 slice  = position;
                    slice[a]  = mn[a] + float(i) * step_a;
                    Candidate c;
                    c.d    = slice[a];
                    c.axis = a;
                    candidates.push_back(c);
                }
            }
        }
        for (Candidate& c : candidates) {
            c.evaluate(begin, end, mapping.data(), bounds, cone_weight, set);
        }
        float    min_cost = candidates[0].cost;
        uint32_t sc       = 0;
        for (uint32_t i = 1; i < candidates.size(); ++i) {
            if (candidates[i].cost < min_cost) {
                sc       = i;
                min_cost = candidates[i].cost;
            }
        }
        return candidates[sc];
    }

    // Tree_builder::assign (light_tree_builder.cpp:737-758)
    uint32_t assign(uint32_t node_id, uint32_t begin, uint32_t end, Box const& bounds,
                    Cone const& cone, float total_power) {
        for (uint32_t i = begin; i < end; ++i) {
            orders[mapping[i]] = order++;
        }
        Build_node& node = nodes[node_id];
        node.bounds      = bounds;
        node.cone        = cone;
        node.power       = total_power;
        node.var         = variance(mapping.data(), begin, end, set);
        node.middle      = 0;
        node.children    = begin;
        node.num_lights  = end - begin;
        node.two_sided   = false;
        return end;
    }

    // Tree_builder::split of a part (light_tree_builder.cpp:661-735)
    uint32_t split_part(uint32_t node_id, uint32_t begin, uint32_t end, Box const& bounds,
                        Cone const& cone, float total_power) {
        uint32_t const len = end - begin;
        if (len <= max_primitives) {
            return assign(node_id, begin, end, bounds, cone, total_power);
        }
        uint32_t const child0      = current;
        float const    cone_weight = cone_cost(cone.v[3]);
        Candidate      sc          = evaluate_splits(begin, end, bounds, cone_weight);
        if (sc.exhausted) {
            return assign(node_id, begin, end, bounds, cone, total_power);
        }
        uint32_t const split_node = uint32_t(
            std::partition(mapping.data() + begin, mapping.data() + end,
                           [&](uint32_t l) { return sc.behind(set.boxes[l].hi); }) -
            mapping.data());
        current += 2;
        uint32_t const c0_end = split_part(child0, begin, split_node, sc.box0, sc.cone0, sc.power0);
        uint32_t const c1_end = split_part(child0 + 1, split_node, end, sc.box1, sc.cone1, sc.power1);
        Build_node& node = nodes[node_id];
        node.bounds      = bounds;
        node.cone        = cone;
        node.power       = total_power;
        node.var         = variance(mapping.data(), begin, end, set);
        node.middle      = c0_end;
        node.children    = child0;
        node.num_lights  = len;
        node.two_sided   = false;
        return c1_end;
    }

    uint32_t split(uint32_t node_id, uint32_t begin, uint32_t end, Box const& bounds,
                   Cone const& cone, bool two_sided, float total_power) {
        uint32_t const len = end - begin;
        if (len <= 4) {
            bool node_two_sided = false;
            for (uint32_t i = begin; i < end; ++i) {
                orders[mapping[i]] = order++;
                node_two_sided |= 0 != set.two_sided[mapping[i]];
            }
            Build_node& node = nodes[node_id];
            node.bounds      = bounds;
            node.cone        = cone;
            node.power       = total_power;
            node.var         = variance(mapping.data(), begin, end, set);
            node.middle      = 0;
            node.children    = begin;
            node.num_lights  = len;
            node.two_sided   = node_two_sided;
            return begin + len;
        }
        uint32_t const child0 = current;
        current += 2;
        float const cone_weight = cone_cost(cone.v[3]);
        Candidate   sc          = evaluate_splits(begin, end, bounds, cone_weight);
        uint32_t    split_node  = uint32_t(
            std::partition(mapping.data() + begin, mapping.data() + end,
                           [&](uint32_t l) { return sc.behind(set.boxes[l].hi); }) -
            mapping.data());
        if (split_node == begin || split_node == end) {
            // every candidate leaves a side empty (the reference asserts this
            // never happens and recurses forever when it does): split at the
            // median of the box centers along the longest axis instead
            F3 const  e    = bounds.extent();
            int const axis = e[0] >= e[1] ? (e[0] >= e[2] ? 0 : 2) : (e[1] >= e[2] ? 1 : 2);
            std::sort(mapping.data() + begin, mapping.data() + end, [&](uint32_t x, uint32_t y) {
                float const cx = set.boxes[x].lo[axis] + set.boxes[x].hi[axis];
                float const cy = set.boxes[y].lo[axis] + set.boxes[y].hi[axis];
                return cx < cy || (cx == cy && x < y);
            });
            split_node = begin + (end - begin) / 2;
            Box  b0, b1;
            Cone c0, c1;
            bool t0 = false, t1 = false;
            float p0 = 0.f, p1 = 0.f;
            for (uint32_t i = begin; i < end; ++i) {
                uint32_t const l = mapping[i];
                if (i < split_node) {
                    b0.merge(set.boxes[l]), c0 = merge(c0, set.cones[l]), t0 |= 0 != set.two_sided[l], p0 += set.powers[l];
                } else {
                    b1.merge(set.boxes[l]), c1 = merge(c1, set.cones[l]), t1 |= 0 != set.two_sided[l], p1 += set.powers[l];
                }
            }
            sc.box0 = b0, sc.box1 = b1, sc.cone0 = c0, sc.cone1 = c1, sc.two0 = t0, sc.two1 = t1;
            sc.power0 = p0, sc.power1 = p1;
        }
        uint32_t const c0_end = split(child0, begin, split_node, sc.box0, sc.cone0, sc.two0, sc.power0);
        uint32_t const c1_end = split(child0 + 1, split_node, end, sc.box1, sc.cone1, sc.two1, sc.power1);
        Build_node& node = nodes[node_id];
        node.bounds      = bounds;
        node.cone        = cone;
        node.power       = total_power;
        node.var         = variance(mapping.data(), begin, end, set);
        node.middle      = c0_end;
        node.children    = child0;
        node.num_lights  = len;
        node.two_sided   = two_sided;
        return c1_end;
    }
};

void load_lights(float const* box_min, float const* box_max, float const* cones,
                 float const* powers, uint8_t const* two_sided, uint32_t num_lights, Lights& set) {
    set.boxes.resize(num_lights);
    set.cones.resize(num_lights);
    set.powers.assign(powers, powers + num_lights);
    if (two_sided) {
        set.two_sided.assign(two_sided, two_sided + num_lights);
    } else {
        set.two_sided.assign(num_lights, 0);
    }
    for (uint32_t l = 0; l < num_lights; ++l) {
        for (int k = 0; k < 3; ++k) {
            set.boxes[l].lo[k] = box_min[3 * l + k];
            set.boxes[l].hi[k] = box_max[3 * l + k];
        }
        for (int k = 0; k < 4; ++k) {
            set.cones[l].v[k] = cones[4 * l + k];
        }
    }
}

void serialize(Builder const& b, HostLightTree& out) {
    // serialize, light_tree_builder.cpp:712-733
    out.nodes.resize(b.current);
    out.middles.resize(b.current);
    for (uint32_t i = 0; i < b.current; ++i) {
        Build_node const& s = b.nodes[i];
        Light_node&       d = out.nodes[i];
        F3 const          p = s.bounds.position();
        F3 const          e = s.bounds.extent();
        d.center[0]         = p[0];
        d.center[1]         = p[1];
        d.center[2]         = p[2];
        d.radius            = 0.5f * std::sqrt(dot3(e.v, e.v));
        for (int k = 0; k < 4; ++k) {
            d.cone[k] = s.cone.v[k];
        }
        d.power      = s.power;
        d.variance   = s.var;
        d.bits       = (s.middle > 0 ? 1u : 0u) | (s.two_sided ? 2u : 0u) | (s.children << 2);
        d.num_lights = s.num_lights;
        out.middles[i] = s.middle;
    }
}


}  // namespace

void build_light_tree(float const* box_min, float const* box_max, float const* cones,
                      float const* powers, uint8_t const* two_sided, uint32_t num_lights,
                      HostLightTree& out) {
    out = HostLightTree();
    if (0 == num_lights) {
        return;
    }
    Lights set;
    load_lights(box_min, box_max, cones, powers, two_sided, num_lights, set);

    out.light_mapping.resize(num_lights);
    out.light_orders.resize(num_lights);
    for (uint32_t l = 0; l < num_lights; ++l) {
        out.light_mapping[l] = l;
    }

    Builder b{set, out.light_mapping, out.light_orders, {}, {}, 1, 0};
    b.nodes.resize(2 * size_t(num_lights) - 1);

    Box   bounds;
    Cone  cone;
    bool  ts    = false;
    float total = 0.f;
    for (uint32_t l = 0; l < num_lights; ++l) {
        bounds.merge(set.boxes[l]);
        cone = merge(cone, set.cones[l]);
        ts |= 0 != set.two_sided[l];
        total += set.powers[l];
    }
    b.split(0, 0, num_lights, bounds, cone, ts, total);
    serialize(b, out);
}

void build_part_tree(float const* box_min, float const* box_max, float const* cones,
                     float const* powers, uint32_t num_lights, float const part_cone[4],
                     HostLightTree& out) {
    out = HostLightTree();
    if (0 == num_lights) {
        return;
    }
    Lights set;
    load_lights(box_min, box_max, cones, powers, nullptr, num_lights, set);
    out.light_mapping.resize(num_lights);
    out.light_orders.resize(num_lights);
    for (uint32_t l = 0; l < num_lights; ++l) {
        out.light_mapping[l] = l;
    }
    Builder b{set, out.light_mapping, out.light_orders, {}, {}, 1, 0};
    b.part           = true;
    b.max_primitives = std::max(num_lights / 64, 4u);
    b.nodes.resize(2 * size_t(num_lights) - 1);

    // Part::init (triangle_mesh.cpp:177-239): the box of the lights with
    // power, the part's cone, the total power
    Box   bounds;
    float total = 0.f;
    for (uint32_t l = 0; l < num_lights; ++l) {
        if (set.powers[l] > 0.f) {
            bounds.merge(set.boxes[l]);
            total += set.powers[l];
        }
    }
    Cone cone;
    for (int k = 0; k < 4; ++k) {
        cone.v[k] = part_cone[k];
    }
    b.split_part(0, 0, num_lights, bounds, cone, total);
    serialize(b, out);

    // leaves of more than 4 lights sample Distribution_1D over the powers
    // (light_tree_builder.cpp:760-795, distribution_1d.cpp:42-82)
    out.cdf_offset.assign(out.nodes.size(), 0xFFFFFFFFu);
    for (uint32_t i = 0; i < out.nodes.size(); ++i) {
        Light_node const& nd = out.nodes[i];
        if ((nd.bits & 1u) || nd.num_lights <= 4) {
            continue;
        }
        uint32_t const first = nd.bits >> 2;
        uint32_t const num   = nd.num_lights;
        float          integral = 0.f;
        for (uint32_t t = 0; t < num; ++t) {
            integral += set.powers[out.light_mapping[first + t]];
        }
        out.cdf_offset[i] = uint32_t(out.cdfs.size());
        if (0.f == integral) {
            out.cdfs.push_back(1.f);  // size 1: {1, 1}
            out.cdfs.push_back(1.f);
            continue;
        }
        float const ii = 1.f / integral;
        out.cdfs.push_back(0.f);
        for (uint32_t t = 1; t < num; ++t) {
            out.cdfs.push_back(std::fma(set.powers[out.light_mapping[first + t - 1]], ii, out.cdfs.back()));
        }
        out.cdfs.push_back(1.f);
    }
}

}  // namespace mi
