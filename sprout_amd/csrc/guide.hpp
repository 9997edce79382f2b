// Path-guiding network: parameter layout, device-side mixture of von
// Mises-Fisher lobes (sampling, density), and the host interface of the
// BF16 MFMA inference and training kernels (guide.hip).
//
// Network: 9 inputs (position in the scene box, normal and outgoing
// direction mapped to [0,1]) -> 12-frequency sin/cos encoding, sin(pi f x)
// and cos(pi f x) for 12 multipliers f per input (216 features, zero-padded
// to 224; by default f = 2^(j/2) for the position and 2^(j/4) for normal and
// direction, j = 0..11) -> 64 -> 64 -> 64 (ReLU) -> 32 = 8 lobes x
// (a, b, c, d): theta = pi sigmoid(a), phi = 2 pi sigmoid(b),
// kappa = exp(clamp(c, -4, 8)), weight = softmax(d).
#pragma once

#include "bsdf.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mi {

namespace guide {

inline constexpr int Num_in    = 9;
inline constexpr int Num_freq  = 12;
inline constexpr int Num_feat  = Num_in * Num_freq * 2;  // 216
inline constexpr int Feat_pad  = 224;
inline constexpr int Hidden    = 64;
inline constexpr int Num_lobes = 8;
inline constexpr int Num_out   = 4 * Num_lobes;

// fp32 parameters: per layer W [in][out] row-major, then b [out]
inline constexpr int Off_w1     = 0;
inline constexpr int Off_b1     = Off_w1 + Feat_pad * Hidden;
inline constexpr int Off_w2     = Off_b1 + Hidden;
inline constexpr int Off_b2     = Off_w2 + Hidden * Hidden;
inline constexpr int Off_w3     = Off_b2 + Hidden;
inline constexpr int Off_b3     = Off_w3 + Hidden * Hidden;
inline constexpr int Off_w4     = Off_b3 + Hidden;
inline constexpr int Off_b4     = Off_w4 + Hidden * Num_out;
inline constexpr int Num_params = Off_b4 + Num_out;  // 24800

// bf16 weights transposed, W^T [out][in]
inline constexpr int T_w1  = 0;
inline constexpr int T_w2  = T_w1 + Hidden * Feat_pad;
inline constexpr int T_w3  = T_w2 + Hidden * Hidden;
inline constexpr int T_w4  = T_w3 + Hidden * Hidden;
inline constexpr int Num_t = T_w4 + Num_out * Hidden;  // 24576

inline constexpr float Kappa_lo = -4.f;
inline constexpr float Kappa_hi = 8.f;

}  // namespace guide

struct GuideNet {
    float*    params  = nullptr;  // fp32 master parameters
    float*    adam_m  = nullptr;
    float*    adam_v  = nullptr;
    float*    grad    = nullptr;  // reduced gradient of the last batch
    float*    partial = nullptr;  // per-block gradient partial sums
    float*    wsum    = nullptr;  // per-block partial sums of the batch weights
    float*    freq    = nullptr;  // encoding multipliers, Num_in x Num_freq
    uint16_t* wt      = nullptr;  // bf16 W^T of every layer
    uint32_t* window  = nullptr;  // {first record, count, Adam step} of the current batch
    uint32_t* step_dev     = nullptr;  // Adam steps taken (device word)
    uint32_t* scratch_step = nullptr;  // step word of gradient-only windows
    uint32_t  partial_blocks = 0;

    void release();
};

int32_t guide_create(uint64_t seed, uint32_t num_cus, GuideNet& out);
int32_t guide_set_params(GuideNet& net, float const* host_params);
// Encoding multipliers (host, Num_in x Num_freq, finite, |f| <= 4096).
int32_t guide_set_frequencies(GuideNet& net, float const* host_freq);
// The default multipliers (Num_in x Num_freq) into out.
void guide_default_frequencies(float* out);
int32_t guide_get_params(GuideNet const& net, float* host_params);
// x: n x 9 inputs in [0,1]; out: n x 32 raw outputs (device arrays)
int32_t guide_forward(GuideNet const& net, float const* x, uint32_t n, float* out,
                      hipStream_t stream);
// Same, over a query count that lives in device memory (at most max_n).
int32_t guide_forward_device_count(GuideNet const& net, float const* x, uint32_t const* n,
                                   uint32_t max_n, float* out, hipStream_t stream);
// Gradient of the batch loss into net.grad (device), no update.
int32_t guide_gradients(GuideNet& net, float const* x, float const* dirs, float const* weights,
                        uint32_t n, hipStream_t stream);
// Gradient + one Adam step.
int32_t guide_adam_step(GuideNet& net, float const* x, float const* dirs, float const* weights,
                        uint32_t n, float lr, hipStream_t stream);
// `steps` Adam steps over the first min(*num_records, capacity) records (a
// count in device memory), in slices of `batch` (0: count / steps); trained
// (device, optional) sums the records used.  Kernels only: capturable.
int32_t guide_train_records(GuideNet& net, float const* x, float const* dirs, float const* weights,
                            uint32_t const* num_records, uint32_t capacity, uint32_t steps,
                            uint32_t batch, float lr, unsigned long long* trained,
                            hipStream_t stream);

namespace dev {

// A mixture as Vmf_mixture::store writes it: per lobe mu (3), kappa, the
// amplitude (weight * kappa / (2 pi (1 - exp(-2 kappa)))) and the weight.
inline constexpr int Vmf_stride = 6;

// Its density: the sum over the lobes of amplitude * exp(kappa (mu.w - 1)).
__device__ __forceinline__ float vmf_mixture_pdf(float const* lobes, V3 w) {
    float q = 0.f;
#pragma unroll
    for (int k = 0; k < guide::Num_lobes; ++k) {
        float const* l = lobes + Vmf_stride * k;
        float const  c = l[0] * w.x + l[1] * w.y + l[2] * w.z;
        q += l[4] * expf(l[3] * (c - 1.f));
    }
    return q;
}

// A direction from it: the lobe by u0 (the last unless the running sum of
// the weights passes u0 earlier), then w.mu = 1 + log(u1 + (1 - u1)
// e^(-2 kappa)) / kappa and the azimuth 2 pi u2.
__device__ inline V3 vmf_mixture_sample(float const* lobes, float u0, float u1, float u2) {
    int   lobe = guide::Num_lobes - 1;
    float acc  = 0.f;
    for (int k = 0; k < guide::Num_lobes - 1; ++k) {
        acc += lobes[Vmf_stride * k + 5];
        if (u0 < acc) {
            lobe = k;
            break;
        }
    }
    float const* l   = lobes + Vmf_stride * lobe;
    V3 const     m   = v3(l[0], l[1], l[2]);
    float const  kap = l[3];
    float const  wz  = fminf(1.f, fmaxf(-1.f, 1.f + logf(u1 + (1.f - u1) * expf(-2.f * kap)) / kap));
    float const  r   = sqrtf(fmaxf(0.f, 1.f - wz * wz));
    float        sp, cp;
    sincosf(2.f * Pi * u2, &sp, &cp);
    Layer const fr = frame_from_normal(m);
    return normalize(fr.tangent_to_world(v3(r * cp, r * sp, wz)));
}

// The mixture of one query, ready for sampling and density evaluation.
struct Vmf_mixture {
    static constexpr int Stored = Vmf_stride * guide::Num_lobes;  // floats store writes

    float mu[guide::Num_lobes][3];
    float kappa[guide::Num_lobes];
    float amp[guide::Num_lobes];     // weight * kappa / (2 pi (1 - exp(-2 kappa)))
    float weight[guide::Num_lobes];

    __device__ void set(float const* raw) {
        float dmax = raw[3];
#pragma unroll
        for (int k = 1; k < guide::Num_lobes; ++k) {
            dmax = fmaxf(dmax, raw[4 * k + 3]);
        }
        float wsum = 0.f;
#pragma unroll
        for (int k = 0; k < guide::Num_lobes; ++k) {
            float const* r     = raw + 4 * k;
            float const  sa    = 1.f / (1.f + expf(-r[0]));
            float const  sb    = 1.f / (1.f + expf(-r[1]));
            float const  theta = Pi * sa;
            float const  phi   = 2.f * Pi * sb;
            float        st, ct, sp, cp;
            sincosf(theta, &st, &ct);
            sincosf(phi, &sp, &cp);
            mu[k][0]        = st * cp;
            mu[k][1]        = st * sp;
            mu[k][2]        = ct;
            float const kap = expf(fminf(fmaxf(r[2], guide::Kappa_lo), guide::Kappa_hi));
            kappa[k]        = kap;
            weight[k]       = expf(r[3] - dmax);
            wsum += weight[k];
        }
#pragma unroll
        for (int k = 0; k < guide::Num_lobes; ++k) {
            weight[k] /= wsum;
            amp[k] = weight[k] * (kappa[k] / (2.f * Pi * -expm1f(-2.f * kappa[k])));
        }
    }

    // Stored floats per lobe: mu (3), kappa, amplitude, weight;
    // vmf_mixture_pdf of them equals pdf() bit for bit, and
    // vmf_mixture_sample samples them.
    __device__ void store(float* out) const {
#pragma unroll
        for (int k = 0; k < guide::Num_lobes; ++k) {
            out[Vmf_stride * k + 0] = mu[k][0];
            out[Vmf_stride * k + 1] = mu[k][1];
            out[Vmf_stride * k + 2] = mu[k][2];
            out[Vmf_stride * k + 3] = kappa[k];
            out[Vmf_stride * k + 4] = amp[k];
            out[Vmf_stride * k + 5] = weight[k];
        }
    }

    __device__ float pdf(V3 w) const {
        float q = 0.f;
#pragma unroll
        for (int k = 0; k < guide::Num_lobes; ++k) {
            float const c = mu[k][0] * w.x + mu[k][1] * w.y + mu[k][2] * w.z;
            q += amp[k] * expf(kappa[k] * (c - 1.f));
        }
        return q;
    }
};

}  // namespace dev

}  // namespace mi
