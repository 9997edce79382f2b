// Host interface of the device light-tree builder (light_build.hip).
#pragma once

#include "common.hpp"
#include "light_build.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mi {

// The trees to build: num_roots trees over consecutive light ranges
// first[r] .. first[r + 1] (host array of num_roots + 1), every range non
// empty.  Light arrays are device arrays indexed by global light.
struct Light_forest {
    lb::Light_set   set;
    uint32_t const* first      = nullptr;  // host
    float const*    part_cones = nullptr;  // device, 4 per root (parts)
    uint32_t        num_roots  = 0;
    uint32_t        num_lights = 0;
    bool            part       = false;  // Primitive_trees (else scene trees)
};

// Device outputs.  Root r's nodes and middles start at 2 first[r] - r (room
// for 2 n_r - 1), its cdfs at 2 first[r] (room for 2 n_r); mapping (position
// -> light) and orders (light -> position) are local to the root.
struct Light_forest_out {
    mi_light_node* nodes      = nullptr;
    uint32_t*      middles    = nullptr;
    uint32_t*      mapping    = nullptr;
    uint32_t*      orders     = nullptr;
    uint32_t*      cdf_offset = nullptr;  // parts: per node, relative to the root's cdfs
    float*         cdfs       = nullptr;
    uint32_t*      num_cdfs   = nullptr;  // device, per root
    uint32_t*      num_nodes_host = nullptr;  // host, per root (optional)
};

struct Light_forest_scratch {
    uint32_t* mapping    = nullptr;
    void*     nodes      = nullptr;
    uint32_t* counters   = nullptr;
    uint32_t* cand_node  = nullptr;
    uint32_t* cand_k     = nullptr;
    float*    cand_cost  = nullptr;
    uint32_t* cand_start = nullptr;
    uint32_t* cand_count = nullptr;
    uint32_t* node_base  = nullptr;
    uint32_t* light_root = nullptr;
    uint32_t* cdf_base   = nullptr;
    uint32_t* num_nodes  = nullptr;
    uint32_t* first      = nullptr;
    uint32_t* counters_host = nullptr;  // pinned
    uint32_t  light_capacity = 0;
    uint32_t  root_capacity  = 0;

    int32_t reserve(uint32_t lights, uint32_t roots);
    void    release();
};

// Builds the forest on `stream` and waits for it (a level at a time).
// MI_ERR_UNSPLITTABLE: a scene-tree node no candidate splits (the host
// builder's median fallback); nothing else is wrong with the inputs.
int32_t build_light_forest(Light_forest const& in, Light_forest_scratch& scratch,
                           Light_forest_out& out, hipStream_t stream);

}  // namespace mi
