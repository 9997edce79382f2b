// C entry points over the reference's sensor filters (Gaussian, Blackman,
// Mitchell; core/rendering/sensor/filter/), built by ref.mk into
// oracle/_ref.  Test infrastructure only.

#include "rendering/sensor/filter/sensor_blackman.hpp"
#include "rendering/sensor/filter/sensor_gaussian.hpp"
#include "rendering/sensor/filter/sensor_mitchell.hpp"

#include <cstdint>

extern "C" {

// kind 1 Gaussian(radius, alpha), 2 Blackman(radius), 3 Mitchell(radius, b, c):
// out[i] = filter.evaluate(d[i])
void sref_filter_eval(int32_t kind, float radius, float alpha, float b, float c, float const* d,
                      uint32_t n, float* out) {
    using namespace rendering::sensor::filter;
    if (1 == kind) {
        Gaussian const f(radius, alpha);
        for (uint32_t i = 0; i < n; ++i) {
            out[i] = f.evaluate(d[i]);
        }
    } else if (2 == kind) {
        Blackman const f(radius);
        for (uint32_t i = 0; i < n; ++i) {
            out[i] = f.evaluate(d[i]);
        }
    } else {
        Mitchell const f(radius, b, c);
        for (uint32_t i = 0; i < n; ++i) {
            out[i] = f.evaluate(d[i]);
        }
    }
}
}
