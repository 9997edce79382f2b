// Shared host-side definitions of the sprout_gpu library.
#pragma once

#include "sprout_gpu.h"

#include <cstdint>

namespace mi {

// Records the message as this thread's last error and returns status.
int32_t fail(int32_t status, char const* fmt, ...) __attribute__((format(printf, 2, 3)));

void clear_error();

// Stack entries a traversal needs for a tree of the given depth: every
// pending entry is the far child of an ancestor of the current node, plus
// the reference loop's sentinel (node_stack.hpp:8-28, triangle_bvh_tree.cpp:47).
inline uint32_t stack_entries(uint32_t max_depth) {
    return max_depth + 2;
}

// sprout's Node_stack holds 127 entries (core/scene/shape/node_stack.hpp:24).
inline constexpr uint32_t Reference_stack_size = 127;

}  // namespace mi
