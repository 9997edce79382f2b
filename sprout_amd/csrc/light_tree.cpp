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

#include "light_build.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace mi {

namespace {

using lb::Box;
using lb::Candidate;
using lb::Cone;

struct Lights {
    std::vector<float>   lo, hi;  // 3 per light
    std::vector<float>   cones;   // 4 per light
    std::vector<float>   powers;
    std::vector<uint8_t> two_sided;

    lb::Light_set view() const {
        return {lo.data(), hi.data(), cones.data(), powers.data(), two_sided.data()};
    }
    float const* hi_of(uint32_t l) const { return hi.data() + 3 * size_t(l); }
    float const* lo_of(uint32_t l) const { return lo.data() + 3 * size_t(l); }
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

    // evaluate_splits (light_tree_builder.cpp:346-420): the cheapest
    // candidate, the first of equal ones; none (a node whose lights sit at
    // one point, too many for the sweep) gives an exhausted candidate
    Candidate evaluate_splits(uint32_t begin, uint32_t end, Box const& bounds, float cone_weight) {
        uint32_t const  len  = end - begin;
        lb::Light_set const view = set.view();
        bool const      sweep = len < (part ? lb::Part_sweep_threshold : lb::Scene_sweep_threshold);
        uint32_t const  n     = sweep ? 3 * len : lb::grid_candidates(bounds);
        candidates.resize(n);
        for (uint32_t k = 0; k < n; ++k) {
            Candidate& c = candidates[k];
            c            = Candidate();
            if (sweep) {
                lb::sweep_candidate(mapping.data(), begin, k, view, c);
            } else {
                lb::grid_candidate(bounds, k, c);
            }
            c.evaluate(begin, end, mapping.data(), bounds, cone_weight, view);
        }
        if (0 == n) {
            Candidate c;
            c.exhausted = true;
            c.d         = -3.40282347e+38f;  // every light on the far side
            return c;
        }
        float    min_cost = candidates[0].cost;
        uint32_t sc       = 0;
        for (uint32_t i = 1; i < n; ++i) {
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
        node.var         = lb::variance(mapping.data(), begin, end, set.view());
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
        float const    cone_weight = lb::cone_cost(cone.v[3]);
        Candidate      sc          = evaluate_splits(begin, end, bounds, cone_weight);
        if (sc.exhausted) {
            return assign(node_id, begin, end, bounds, cone, total_power);
        }
        uint32_t const split_node = uint32_t(
            std::partition(mapping.data() + begin, mapping.data() + end,
                           [&](uint32_t l) { return sc.behind(set.hi_of(l)); }) -
            mapping.data());
        current += 2;
        uint32_t const c0_end = split_part(child0, begin, split_node, sc.box0, sc.cone0, sc.power0);
        uint32_t const c1_end = split_part(child0 + 1, split_node, end, sc.box1, sc.cone1, sc.power1);
        Build_node& node = nodes[node_id];
        node.bounds      = bounds;
        node.cone        = cone;
        node.power       = total_power;
        node.var         = lb::variance(mapping.data(), begin, end, set.view());
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
                node_two_sided |= set.view().two(mapping[i]);
            }
            Build_node& node = nodes[node_id];
            node.bounds      = bounds;
            node.cone        = cone;
            node.power       = total_power;
            node.var         = lb::variance(mapping.data(), begin, end, set.view());
            node.middle      = 0;
            node.children    = begin;
            node.num_lights  = len;
            node.two_sided   = node_two_sided;
            return begin + len;
        }
        uint32_t const child0 = current;
        current += 2;
        float const cone_weight = lb::cone_cost(cone.v[3]);
        Candidate   sc          = evaluate_splits(begin, end, bounds, cone_weight);
        uint32_t    split_node  = uint32_t(
            std::partition(mapping.data() + begin, mapping.data() + end,
                           [&](uint32_t l) { return sc.behind(set.hi_of(l)); }) -
            mapping.data());
        if (split_node == begin || split_node == end) {
            // every candidate leaves a side empty (the reference asserts this
            // never happens and recurses forever when it does): split at the
            // median of the box centers along the longest axis instead
            float const e0 = bounds.extent(0), e1 = bounds.extent(1), e2 = bounds.extent(2);
            int const   axis = e0 >= e1 ? (e0 >= e2 ? 0 : 2) : (e1 >= e2 ? 1 : 2);
            std::sort(mapping.data() + begin, mapping.data() + end, [&](uint32_t x, uint32_t y) {
                float const cx = set.lo_of(x)[axis] + set.hi_of(x)[axis];
                float const cy = set.lo_of(y)[axis] + set.hi_of(y)[axis];
                return cx < cy || (cx == cy && x < y);
            });
            split_node = begin + (end - begin) / 2;
            Box  b0, b1;
            Cone c0, c1;
            bool t0 = false, t1 = false;
            float p0 = 0.f, p1 = 0.f;
            for (uint32_t i = begin; i < end; ++i) {
                uint32_t const l = mapping[i];
                lb::Light_set const view = set.view();
                if (i < split_node) {
                    b0.merge(set.lo_of(l), set.hi_of(l)), c0 = lb::merge(c0, view.cone(l)), t0 |= view.two(l), p0 += set.powers[l];
                } else {
                    b1.merge(set.lo_of(l), set.hi_of(l)), c1 = lb::merge(c1, view.cone(l)), t1 |= view.two(l), p1 += set.powers[l];
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
        node.var         = lb::variance(mapping.data(), begin, end, set.view());
        node.middle      = c0_end;
        node.children    = child0;
        node.num_lights  = len;
        node.two_sided   = two_sided;
        return c1_end;
    }
};

void load_lights(float const* box_min, float const* box_max, float const* cones,
                 float const* powers, uint8_t const* two_sided, uint32_t num_lights, Lights& set) {
    set.lo.assign(box_min, box_min + 3 * size_t(num_lights));
    set.hi.assign(box_max, box_max + 3 * size_t(num_lights));
    set.cones.assign(cones, cones + 4 * size_t(num_lights));
    set.powers.assign(powers, powers + num_lights);
    if (two_sided) {
        set.two_sided.assign(two_sided, two_sided + num_lights);
    } else {
        set.two_sided.assign(num_lights, 0);
    }
}

void serialize(Builder const& b, HostLightTree& out) {
    // serialize, light_tree_builder.cpp:712-733
    out.nodes.resize(b.current);
    out.middles.resize(b.current);
    for (uint32_t i = 0; i < b.current; ++i) {
        Build_node const& s = b.nodes[i];
        Light_node&       d = out.nodes[i];
        float const       e[3] = {s.bounds.extent(0), s.bounds.extent(1), s.bounds.extent(2)};
        d.center[0]            = s.bounds.position(0);
        d.center[1]            = s.bounds.position(1);
        d.center[2]            = s.bounds.position(2);
        d.radius               = 0.5f * std::sqrt(lb::dot3(e, e));
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
    lb::Light_set const view = set.view();
    for (uint32_t l = 0; l < num_lights; ++l) {
        bounds.merge(set.lo_of(l), set.hi_of(l));
        cone = lb::merge(cone, view.cone(l));
        ts |= view.two(l);
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
            bounds.merge(set.lo_of(l), set.hi_of(l));
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
