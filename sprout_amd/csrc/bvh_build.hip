#include "bvh_build.hpp"

namespace mi {

int32_t build_bvh(float const*, uint32_t, uint32_t, uint32_t const*, uint32_t,
                  mi_bvh_settings const&, hipStream_t, DeviceBvh&) {
    return fail(MI_ERR_ARGS, "mi_bvh_build: not available yet");
}

}  // namespace mi
