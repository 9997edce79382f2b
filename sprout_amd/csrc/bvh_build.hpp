// GPU construction of sprout's SBVH (host interface).
#pragma once

#include "common.hpp"
#include "trace.hpp"

#include <hip/hip_runtime.h>

namespace mi {

// Builds the tree scene::shape::triangle::bvh::Builder_SAH builds
// (triangle_bvh_builder_sah.cpp:21-67) from a device-resident mesh, on the
// device, and lays it out depth-first as serialize() does.  Synchronous.
int32_t build_bvh(float const* positions, uint32_t num_vertices, uint32_t position_stride,
                  uint32_t const* indices, uint32_t num_triangles, mi_bvh_settings const& settings,
                  hipStream_t stream, DeviceBvh& out);

}  // namespace mi
