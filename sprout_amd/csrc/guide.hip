// BF16 MFMA inference and training kernels of the path-guiding network.
//
// Matrix products run on gfx950's v_mfma_f32_16x16x32_bf16 (64-wide wave,
// FP32 accumulation).  Operand lanes, for lane l, r = l % 16, q = l / 16:
//   A (16 x 16, rows i, depth k): A[r][4q + t], t = 0..3  (x32: A[r][8q + t], t = 0..7)
//   B (16 x 16, depth k, cols j): B[4q + t][r]            (x32: B[8q + t][r])
//   D (16 x 16):                   D[4q + t][r]
// Weights live in LDS as bf16 W^T [out][in] so a forward B operand is one
// 8-byte read; activations as bf16 [row][feature] so an A operand is one
// 8-byte read.
//
// forward   a wave computes 16 queries: the first layer takes its features
//           (sinpi / cospi of f x) straight from registers, two hidden layers
//           pass through LDS, the output layer stores to memory; weights
//           are loaded once per workgroup and the grid strides over tiles.
// gradient  a workgroup of 4 waves takes 64 samples at a time: each wave runs
//           forward and backward for its 16 (activations and pre-activation
//           gradients kept in LDS), then the 4 waves split the 96 16x16 tiles
//           of the weight gradients (depth = the 64 samples) and keep them in
//           registers across all chunks; partial sums per workgroup are
//           reduced in a fixed order, so gradients are deterministic.
// adam      one thread per parameter; refreshes the bf16 W^T copy.

#include "common.hpp"
#include "guide.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

namespace mi {

namespace {

using namespace guide;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// LDS row strides (elements), padded against bank conflicts; multiples of 4
constexpr int S_feat = Feat_pad + 8;  // 232
constexpr int S_hid  = Hidden + 8;    // 72
constexpr int S_out  = Num_out + 8;   // 40

typedef float  f32x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

// fp32 -> bf16, rounded to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint16_t to_bf16(float f) {
    return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f));
}

// two bf16 in one word, lo in the low half (one v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
    f32x2_t const f = {lo, hi};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f, bf16x2_t));
}

__device__ __forceinline__ float from_bf16(uint16_t h) {
    return __uint_as_float(uint32_t(h) << 16);
}

// Orders this wave's LDS accesses (writes by some lanes, reads by others).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ bf16x4 load4(uint16_t const* p) {
    return *reinterpret_cast<bf16x4 const*>(p);
}



// gfx950's double-depth product: 16 x 16 x 32, each lane supplying eight
// consecutive depth elements.  Lane (r, q) gives A[r][8q + t] and
// B[8q + t][r], t = 0..7; whatever order the unit adds the 32 depth terms
// in, D is the sum over all of them, so the operands only need the same
// depth index at the same lane and slot.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ bf16x8 load8(uint16_t const* p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<uint4 const*>(p));
}

__device__ __forceinline__ bf16x8 gather8(uint16_t const* p, int stride) {
    uint32_t w[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        w[t] = uint32_t(p[2 * t * stride]) | (uint32_t(p[(2 * t + 1) * stride]) << 16);
    }
    return __builtin_bit_cast(bf16x8, make_uint4(w[0], w[1], w[2], w[3]));
}

__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

struct Weights {
    uint16_t w1[Hidden * S_feat];
    uint16_t w2[Hidden * S_hid];
    uint16_t w3[Hidden * S_hid];
    uint16_t w4[Num_out * S_hid];
    float    b[3 * Hidden + Num_out];
};

__device__ void load_weights(Weights& s, uint16_t const* wt, float const* params) {
    for (int i = threadIdx.x; i < Hidden * Feat_pad; i += blockDim.x) {
        s.w1[(i / Feat_pad) * S_feat + i % Feat_pad] = wt[T_w1 + i];
    }
    for (int i = threadIdx.x; i < Hidden * Hidden; i += blockDim.x) {
        s.w2[(i / Hidden) * S_hid + i % Hidden] = wt[T_w2 + i];
        s.w3[(i / Hidden) * S_hid + i % Hidden] = wt[T_w3 + i];
    }
    for (int i = threadIdx.x; i < Num_out * Hidden; i += blockDim.x) {
        s.w4[(i / Hidden) * S_hid + i % Hidden] = wt[T_w4 + i];
    }
    for (int i = threadIdx.x; i < Hidden; i += blockDim.x) {
        s.b[i]              = params[Off_b1 + i];
        s.b[Hidden + i]     = params[Off_b2 + i];
        s.b[2 * Hidden + i] = params[Off_b3 + i];
    }
    for (int i = threadIdx.x; i < Num_out; i += blockDim.x) {
        s.b[3 * Hidden + i] = params[Off_b4 + i];
    }
}

// Features of 16 queries into X [16][S_feat]; row m = query row0 + m, rows
// past n are zero.  x: 9 inputs per query in [0,1]; feature (i, j) is
// sin / cos of pi (freq[i][j] x_i), the product rounded to fp32.
__device__ void features(float const* x, float const* freq, uint32_t row0, uint32_t n, uint16_t* X) {
    int const      lane = threadIdx.x & 63;
    int const      m    = lane & 15;
    int const      g    = lane >> 4;
    uint32_t const row  = row0 + m;
    for (int i = g; i < Num_in; i += 4) {
        float const v = row < n ? x[size_t(row) * Num_in + i] : 0.f;
        for (int j = 0; j < Num_freq; ++j) {
            // sin(pi f) = sin(2 pi (f / 2)): the hardware sine of the
            // fraction of f / 2 turns (f / 2 and its fraction are exact)
            float const f   = freq[i * Num_freq + j] * v;
            float const h   = 0.5f * f;
            float const rev = h - floorf(h);
            float const s   = row < n ? __builtin_amdgcn_sinf(rev) : 0.f;
            float const c   = row < n ? __builtin_amdgcn_cosf(rev) : 0.f;
            // the pair (sin, cos) as one 4-byte store (row and pair offsets are even)
            *reinterpret_cast<uint32_t*>(X + m * S_feat + 24 * i + 2 * j) = pack_bf16(s, c);
        }
    }
    if (3 == g) {
        for (int k = Num_feat; k < Feat_pad; ++k) {
            X[m * S_feat + k] = 0;
        }
    }
}

// Epilogue of a layer: acc[nt][t] is output 16 nt + 4q + t of row r; adds
// the bias and writes, per Mode, ReLU'd bf16 to H (one 8-byte store per
// tile), fp32 to F (Mode_f32), or fp32 rows r < rows to global F (Mode_out;
// float4 stores where F is 16-byte aligned, sf a multiple of 4).
enum { Mode_relu = 0, Mode_f32 = 1, Mode_out = 2 };

template <int N, int Mode>
__device__ __forceinline__ void store_layer(f32x4 const (&acc)[N / 16], float const* bias, int r,
                                            int q, uint16_t* H, int sh, float* F, int sf,
                                            int rows) {
#pragma unroll
    for (int nt = 0; nt < N / 16; ++nt) {
        int const j0 = nt * 16 + 4 * q;
        float     v[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            v[t] = acc[nt][t] + bias[j0 + t];
        }
        if (Mode_relu == Mode) {
            uint2 pk;
            pk.x = pack_bf16(fmaxf(v[0], 0.f), fmaxf(v[1], 0.f));
            pk.y = pack_bf16(fmaxf(v[2], 0.f), fmaxf(v[3], 0.f));
            *reinterpret_cast<uint2*>(H + r * sh + j0) = pk;
        } else if (Mode_f32 == Mode) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                F[r * sf + j0 + t] = v[t];
            }
        } else if (r < rows) {
            float* const o = F + size_t(r) * sf + j0;
            if (0 == (reinterpret_cast<uintptr_t>(o) & 15u)) {
                *reinterpret_cast<float4*>(o) = make_float4(v[0], v[1], v[2], v[3]);
            } else {
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    o[t] = v[t];
                }
            }
        }
    }
}

// out[16][N] = act(A[16][K] W + b), formed as its transpose W^T A^T so that a
// lane's four results are four consecutive outputs of one row: MFMA A
// operand = rows of W^T in LDS (stride sw), B operand = activation rows
// (bf16 in LDS, stride sa), D[4q + t][r] = output 16 nt + 4q + t of row r.
// Double-depth products: one 16-byte operand read per lane and step.
// The depth steps run outermost, so each activation operand is read once and
// the N / 16 accumulation chains interleave.
template <int K, int N, int Mode>
__device__ void layer(uint16_t const* A, int sa, uint16_t const* Wt, int sw, float const* bias,
                      uint16_t* H, int sh, float* F, int sf, int rows = 16) {
    int const lane = threadIdx.x & 63;
    int const r    = lane & 15;
    int const q    = lane >> 4;
    f32x4     acc[N / 16];
#pragma unroll
    for (int nt = 0; nt < N / 16; ++nt) {
        acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    static_assert(K % 32 == 0, "depth in steps of 32");
#pragma unroll
    for (int ks = 0; ks < K / 32; ++ks) {
        bf16x8 const b = load8(A + r * sa + ks * 32 + 8 * q);
#pragma unroll
        for (int nt = 0; nt < N / 16; ++nt) {
            acc[nt] = mfma32(load8(Wt + (nt * 16 + r) * sw + ks * 32 + 8 * q), b, acc[nt]);
        }
    }
    store_layer<N, Mode>(acc, bias, r, q, H, sh, F, sf, rows);
}

// First layer of the forward pass with its features formed in registers: the
// B operand of lane (r, q) at depth step ks is features k0 .. k0 + 7, k0 =
// 32 ks + 8q, of query row0 + r, i.e. (sin, cos) of pi f x_i for i = k0 / 24
// and the multipliers j0 .. j0 + 3 with j0 = (k0 % 24) / 2 (zero past the
// 216 features): the values features() writes.  half_freq holds
// f / 2, and (f / 2) x equals (f x) / 2 for every normal product.  Rows past
// n run on zero inputs; rows do not mix and theirs are never stored.
__device__ void layer1_direct(float const* x, float const* half_freq, uint32_t row0, uint32_t n,
                              uint16_t const* Wt, float const* bias, uint16_t* H) {
    int const      lane = threadIdx.x & 63;
    int const      r    = lane & 15;
    int const      q    = lane >> 4;
    uint32_t const row  = row0 + r;
    bool const     live = row < n;
    f32x4          acc[Hidden / 16];
#pragma unroll
    for (int nt = 0; nt < Hidden / 16; ++nt) {
        acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    static_assert(Feat_pad % 32 == 0 && 24 % 8 == 0, "eight features of one input per lane");
#pragma unroll
    for (int ks = 0; ks < Feat_pad / 32; ++ks) {
        int const k0 = ks * 32 + 8 * q;
        uint4     pk = make_uint4(0u, 0u, 0u, 0u);
        if (k0 < Num_feat) {
            int const   i  = k0 / 24;
            int const   j0 = (k0 % 24) / 2;
            float const v  = live ? x[size_t(row) * Num_in + i] : 0.f;
            uint32_t    w[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                float const h   = half_freq[i * Num_freq + j0 + u] * v;
                float const rev = h - floorf(h);
                w[u] = pack_bf16(__builtin_amdgcn_sinf(rev), __builtin_amdgcn_cosf(rev));
            }
            pk = make_uint4(w[0], w[1], w[2], w[3]);
        }
        bf16x8 const b = __builtin_bit_cast(bf16x8, pk);
#pragma unroll
        for (int nt = 0; nt < Hidden / 16; ++nt) {
            acc[nt] = mfma32(load8(Wt + (nt * 16 + r) * S_feat + k0), b, acc[nt]);
        }
    }
    store_layer<Hidden, Mode_relu>(acc, bias, r, q, H, S_hid, nullptr, 0, 16);
}

constexpr int Fwd_waves = 16;

struct alignas(16) Fwd_lds {
    Weights  w;
    float    half_freq[Num_in * Num_freq];
    uint16_t h[Fwd_waves][2][16 * S_hid];
};
static_assert(sizeof(Fwd_lds) <= 160 * 1024, "LDS of a workgroup");
static_assert(offsetof(Fwd_lds, h) % 16 == 0 && (S_hid * sizeof(uint16_t)) % 8 == 0,
              "8-byte aligned activation rows");

// A wave computes 16 queries per tile: layer 1 from register features, two
// hidden layers through LDS, the output layer stored straight to out.
__global__ void __launch_bounds__(Fwd_waves * 64)
    forward_kernel(uint16_t const* wt, float const* params, float const* freq, float const* x,
                   uint32_t n_host, uint32_t const* n_dev, float* out) {
    __shared__ Fwd_lds s;
    uint32_t const n = n_dev ? min(*n_dev, n_host) : n_host;
    load_weights(s.w, wt, params);
    for (int i = threadIdx.x; i < Num_in * Num_freq; i += blockDim.x) {
        s.half_freq[i] = 0.5f * freq[i];
    }
    __syncthreads();

    int const wave  = threadIdx.x >> 6;
    uint16_t* H0    = s.h[wave][0];
    uint16_t* H1    = s.h[wave][1];
    float const* bb = s.w.b;

    for (uint32_t tile = blockIdx.x * Fwd_waves + wave; uint64_t(tile) * 16 < n;
         tile += gridDim.x * Fwd_waves) {
        uint32_t const row0 = tile * 16;
        layer1_direct(x, s.half_freq, row0, n, s.w.w1, bb, H0);
        wave_sync();
        layer<Hidden, Hidden, Mode_relu>(H0, S_hid, s.w.w2, S_hid, bb + Hidden, H1, S_hid, nullptr, 0);
        wave_sync();
        layer<Hidden, Hidden, Mode_relu>(H1, S_hid, s.w.w3, S_hid, bb + 2 * Hidden, H0, S_hid, nullptr, 0);
        wave_sync();
        layer<Hidden, Num_out, Mode_out>(H0, S_hid, s.w.w4, S_hid, bb + 3 * Hidden, nullptr, 0,
                                         out + size_t(row0) * Num_out, Num_out,
                                         int(min(n - row0, 16u)));
        wave_sync();
    }
}

// ---- training ----

constexpr int Grad_waves = 4;
constexpr int Chunk      = 16 * Grad_waves;  // samples per workgroup step
constexpr int Grad_tiles = (Feat_pad / 16) * (Hidden / 16) + 2 * (Hidden / 16) * (Hidden / 16) +
                           (Hidden / 16) * (Num_out / 16);  // 56 + 32 + 8 = 96
constexpr int Tiles_per_wave = Grad_tiles / Grad_waves;      // 24

struct alignas(16) Grad_lds {
    Weights  w;
    uint16_t x[Chunk * S_feat];
    uint16_t h[3][Chunk * S_hid];   // post-ReLU activations of the hidden layers
    uint16_t dz[3][Chunk * S_hid];  // loss gradient at the hidden pre-activations
    uint16_t dy[Chunk * S_out];     // loss gradient at the outputs
    float    y[Chunk * S_out];      // outputs
};
static_assert(sizeof(Grad_lds) <= 160 * 1024, "LDS of a workgroup");
static_assert(offsetof(Grad_lds, x) % 8 == 0 && offsetof(Grad_lds, dy) % 8 == 0 &&
                  (S_feat * sizeof(uint16_t)) % 4 == 0 && (S_out * sizeof(uint16_t)) % 8 == 0,
              "packed feature and output-gradient stores are aligned");
constexpr int Num_bias = 3 * Hidden + Num_out;  // 224

// Loss gradient at the outputs of lobes q and q + 4 of one sample (row of
// Y), for g = -t / q: the derivative of -t log q (the reduction divides by
// the batch's weight sum).  The sample's four lanes (q = 0..3, lanes r + 16 q)
// hold two lobes each and share the mixture's sums by lane exchange, so
// every lobe is evaluated once; all 64 lanes must call it.
struct Lobe_terms {
    float v, kap, cosk, sa, sb, theta, phi, d;
};

__device__ Lobe_terms lobe_terms(float const* r, float dirx, float diry, float dirz) {
    Lobe_terms t;
    t.sa  = 1.f / (1.f + expf(-r[0]));
    t.sb  = 1.f / (1.f + expf(-r[1]));
    t.theta = dev::Pi * t.sa;
    t.phi   = 2.f * dev::Pi * t.sb;
    float st, ct, sp, cp;
    sincosf(t.theta, &st, &ct);
    sincosf(t.phi, &sp, &cp);
    t.cosk = (st * cp) * dirx + (st * sp) * diry + ct * dirz;
    t.kap  = expf(fminf(fmaxf(r[2], Kappa_lo), Kappa_hi));
    t.v    = t.kap / (2.f * dev::Pi * -expm1f(-2.f * t.kap)) * expf(t.kap * (t.cosk - 1.f));
    t.d    = r[3];
    return t;
}

__device__ void lobe_gradient(Lobe_terms const& t, float const* r, float g, float q, float wk,
                              float dirx, float diry, float dirz, float out[4]) {
    float const resp = wk * t.v;
    // softmax: dq/dd_k = w_k (v_k - q)
    out[3] = g * wk * (t.v - q);
    // kappa: d log v / d kappa = 1/kappa - 2 e^(-2k) / (1 - e^(-2k)) + (cos - 1)
    float const e2  = expf(-2.f * t.kap);
    float const dlk = 1.f / t.kap - 2.f * e2 / (-expm1f(-2.f * t.kap)) + (t.cosk - 1.f);
    bool const  in  = r[2] > Kappa_lo && r[2] < Kappa_hi;
    out[2]          = in ? g * resp * dlk * t.kap : 0.f;
    // direction: d log v / d mu = kappa w
    float st, ct, sp, cp;
    sincosf(t.theta, &st, &ct);
    sincosf(t.phi, &sp, &cp);
    float const dth = t.kap * ((ct * cp) * dirx + (ct * sp) * diry - st * dirz);
    float const dph = t.kap * ((-st * sp) * dirx + (st * cp) * diry);
    out[0]          = g * resp * dth * dev::Pi * t.sa * (1.f - t.sa);
    out[1]          = g * resp * dph * 2.f * dev::Pi * t.sb * (1.f - t.sb);
}

__device__ float sum4(float v) {
    v += __shfl_xor(v, 16);
    return v + __shfl_xor(v, 32);
}

__device__ void output_gradients(float const* y, float dirx, float diry, float dirz, float tw,
                                 int q, float g0[4], float g1[4]) {
    Lobe_terms const a = lobe_terms(y + 4 * q, dirx, diry, dirz);
    Lobe_terms const b = lobe_terms(y + 4 * (q + 4), dirx, diry, dirz);
    float            dmax = fmaxf(a.d, b.d);
    dmax                  = fmaxf(dmax, __shfl_xor(dmax, 16));
    dmax                  = fmaxf(dmax, __shfl_xor(dmax, 32));
    float const ea        = expf(a.d - dmax);
    float const eb        = expf(b.d - dmax);
    float const esum      = sum4(ea + eb);
    float const wa        = ea / esum;
    float const wb        = eb / esum;
    float const qm        = sum4(wa * a.v + wb * b.v);  // mixture density q
    float const g         = -tw / fmaxf(qm, 1e-30f);
    lobe_gradient(a, y + 4 * q, g, qm, wa, dirx, diry, dirz, g0);
    lobe_gradient(b, y + 4 * (q + 4), g, qm, wb, dirx, diry, dirz, g1);
}

// dst[16][N] = (src[16][K] W^T) * (act > 0): backward through one layer,
// W[j][k] = Wt[k][j].  Formed as its transpose like layer(): MFMA A operand
// A[i][k] = W[16 nt + i][k] = Wt[k][16 nt + i], gathered from 4 rows of Wt;
// B operand = rows of src; so a lane holds outputs 16 nt + 4q + t of row r,
// read against the ReLU mask with one 8-byte load and stored with one.
template <int K, int N>
__device__ void back_layer(uint16_t const* src, int ss, uint16_t const* Wt, int sw,
                           uint16_t const* act, int sact, uint16_t* dst, int sd) {
    int const lane = threadIdx.x & 63;
    int const r    = lane & 15;
    int const q    = lane >> 4;
    f32x4     acc[N / 16];
#pragma unroll
    for (int nt = 0; nt < N / 16; ++nt) {
        acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    static_assert(K % 32 == 0, "depth in steps of 32");
#pragma unroll
    for (int ks = 0; ks < K / 32; ++ks) {
        bf16x8 const b = load8(src + r * ss + ks * 32 + 8 * q);
#pragma unroll
        for (int nt = 0; nt < N / 16; ++nt) {
            acc[nt] = mfma32(gather8(Wt + (ks * 32 + 8 * q) * sw + nt * 16 + r, sw), b, acc[nt]);
        }
    }
#pragma unroll
    for (int nt = 0; nt < N / 16; ++nt) {
        int const    j0 = nt * 16 + 4 * q;
        bf16x4 const m  = load4(act + r * sact + j0);
        float        v[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            v[t] = from_bf16(uint16_t(m[t])) > 0.f ? acc[nt][t] : 0.f;
        }
        *reinterpret_cast<uint2*>(dst + r * sd + j0) =
            make_uint2(pack_bf16(v[0], v[1]), pack_bf16(v[2], v[3]));
    }
}

__global__ void __launch_bounds__(Grad_waves * 64)
    gradient_kernel(uint16_t const* wt, float const* params, float const* freq, float const* x,
                    float const* dirs, float const* weights, uint32_t n, uint32_t const* window,
                    float* partial, float* wsum) {
    __shared__ Grad_lds s;
    if (window) {
        // the batch is records [window[0], window[0] + window[1]) of the arrays
        uint32_t const off = window[0];
        n                  = window[1];
        x += 9 * size_t(off);
        dirs += 3 * size_t(off);
        weights += off;
    }
    load_weights(s.w, wt, params);
    __syncthreads();

    // thread b < 224 sums the gradient of bias b over the samples, in order
    float bias_acc = 0.f;

    int const wave = threadIdx.x >> 6;
    int const lane = threadIdx.x & 63;
    int const r    = lane & 15;
    int const q    = lane >> 4;
    float     wacc = 0.f;  // weights of this lane's samples (q == 0), in order

    f32x4 acc[Tiles_per_wave];
#pragma unroll
    for (int k = 0; k < Tiles_per_wave; ++k) {
        acc[k] = f32x4{0.f, 0.f, 0.f, 0.f};
    }

    uint32_t const chunks = (n + Chunk - 1) / Chunk;
    for (uint32_t c = blockIdx.x; c < chunks; c += gridDim.x) {  // block-uniform
        uint32_t const row0 = c * Chunk + wave * 16;
        uint16_t*      X    = s.x + wave * 16 * S_feat;
        uint16_t*      H1   = s.h[0] + wave * 16 * S_hid;
        uint16_t*      H2   = s.h[1] + wave * 16 * S_hid;
        uint16_t*      H3   = s.h[2] + wave * 16 * S_hid;
        uint16_t*      DZ1  = s.dz[0] + wave * 16 * S_hid;
        uint16_t*      DZ2  = s.dz[1] + wave * 16 * S_hid;
        uint16_t*      DZ3  = s.dz[2] + wave * 16 * S_hid;
        uint16_t*      DY   = s.dy + wave * 16 * S_out;
        float*         Y    = s.y + wave * 16 * S_out;

        features(x, freq, row0, n, X);
        wave_sync();
        layer<Feat_pad, Hidden, Mode_relu>(X, S_feat, s.w.w1, S_feat, s.w.b, H1, S_hid, nullptr, 0);
        wave_sync();
        layer<Hidden, Hidden, Mode_relu>(H1, S_hid, s.w.w2, S_hid, s.w.b + Hidden, H2, S_hid, nullptr, 0);
        wave_sync();
        layer<Hidden, Hidden, Mode_relu>(H2, S_hid, s.w.w3, S_hid, s.w.b + 2 * Hidden, H3, S_hid, nullptr, 0);
        wave_sync();
        layer<Hidden, Num_out, Mode_f32>(H3, S_hid, s.w.w4, S_hid, s.w.b + 3 * Hidden, nullptr, 0, Y, S_out);
        wave_sync();

        // output gradients: lane handles sample r, lobes q and q + 4
        {
            uint32_t const row = row0 + r;
            float          g0[4];
            float          g1[4];
            // unnormalized weights: the reduction divides by the batch's weight sum
            float const tw  = row < n ? weights[row] : 0.f;
            float const dx  = row < n ? dirs[3 * size_t(row) + 0] : 0.f;
            float const dyv = row < n ? dirs[3 * size_t(row) + 1] : 0.f;
            float const dz  = row < n ? dirs[3 * size_t(row) + 2] : 1.f;
            if (0 == q) {
                wacc += tw;
            }
            output_gradients(Y + r * S_out, dx, dyv, dz, tw, q, g0, g1);
            wave_sync();  // Y is read by all lanes before DY and DZ1 are written
            // four consecutive outputs per store (offsets are multiples of 4 elements)
            *reinterpret_cast<uint2*>(DY + r * S_out + 4 * q) =
                make_uint2(pack_bf16(g0[0], g0[1]), pack_bf16(g0[2], g0[3]));
            *reinterpret_cast<uint2*>(DY + r * S_out + 4 * (q + 4)) =
                make_uint2(pack_bf16(g1[0], g1[1]), pack_bf16(g1[2], g1[3]));
        }
        wave_sync();
        back_layer<Num_out, Hidden>(DY, S_out, s.w.w4, S_hid, H3, S_hid, DZ3, S_hid);
        wave_sync();
        back_layer<Hidden, Hidden>(DZ3, S_hid, s.w.w3, S_hid, H2, S_hid, DZ2, S_hid);
        wave_sync();
        back_layer<Hidden, Hidden>(DZ2, S_hid, s.w.w2, S_hid, H1, S_hid, DZ1, S_hid);
        __syncthreads();

        // bias gradients: column sums over the chunk's samples
        if (int(threadIdx.x) < Num_bias) {
            int const       b   = threadIdx.x;
            uint16_t const* src = b < 3 * Hidden ? s.dz[b / Hidden] + b % Hidden : s.dy + (b - 3 * Hidden);
            int const       st  = b < 3 * Hidden ? S_hid : S_out;
            for (int k = 0; k < Chunk; ++k) {
                bias_acc += from_bf16(src[k * st]);
            }
        }

        // weight gradients over the chunk's 64 samples: dW = act^T dz
#pragma unroll
        for (int k = 0; k < Tiles_per_wave; ++k) {
            int const tile = wave + Grad_waves * k;
            uint16_t const* act;
            uint16_t const* dz;
            int sa, sd, mt, nt;
            if (tile < 56) {
                act = s.x, sa = S_feat, dz = s.dz[0], sd = S_hid, mt = tile / 4, nt = tile % 4;
            } else if (tile < 72) {
                act = s.h[0], sa = S_hid, dz = s.dz[1], sd = S_hid, mt = (tile - 56) / 4, nt = (tile - 56) % 4;
            } else if (tile < 88) {
                act = s.h[1], sa = S_hid, dz = s.dz[2], sd = S_hid, mt = (tile - 72) / 4, nt = (tile - 72) % 4;
            } else {
                act = s.h[2], sa = S_hid, dz = s.dy, sd = S_out, mt = (tile - 88) / 2, nt = (tile - 88) % 2;
            }
#pragma unroll
            for (int ks = 0; ks < Chunk / 32; ++ks) {
                // A[i][k] = act[sample k][feature i]; B[k][j] = dz[sample k][j]
                bf16x8 const a = gather8(act + (ks * 32 + 8 * q) * sa + mt * 16 + r, sa);
                bf16x8 const b = gather8(dz + (ks * 32 + 8 * q) * sd + nt * 16 + r, sd);
                acc[k]         = mfma32(a, b, acc[k]);
            }
        }
        __syncthreads();
    }

    // the workgroup's weight sum, lanes in a fixed order (reusing s.y)
    s.y[threadIdx.x] = wacc;
    __syncthreads();
    if (0 == threadIdx.x) {
        float t = 0.f;
        for (int k = 0; k < Grad_waves * 64; ++k) {
            t += s.y[k];
        }
        wsum[blockIdx.x] = t;
    }

    // partial sums of this workgroup, in the parameter layout
    float* out = partial + size_t(blockIdx.x) * Num_params;
#pragma unroll
    for (int k = 0; k < Tiles_per_wave; ++k) {
        int const tile = wave + Grad_waves * k;
        int       base, cols, mt, nt;
        if (tile < 56) {
            base = Off_w1, cols = Hidden, mt = tile / 4, nt = tile % 4;
        } else if (tile < 72) {
            base = Off_w2, cols = Hidden, mt = (tile - 56) / 4, nt = (tile - 56) % 4;
        } else if (tile < 88) {
            base = Off_w3, cols = Hidden, mt = (tile - 72) / 4, nt = (tile - 72) % 4;
        } else {
            base = Off_w4, cols = Num_out, mt = (tile - 88) / 2, nt = (tile - 88) % 2;
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            int const i = mt * 16 + 4 * q + t;
            int const j = nt * 16 + r;
            out[base + i * cols + j] = acc[k][t];
        }
    }
    if (int(threadIdx.x) < Num_bias) {
        int const b       = threadIdx.x;
        int const offs[4] = {Off_b1, Off_b2, Off_b3, Off_b4};
        out[offs[b < 3 * Hidden ? b / Hidden : 3] + (b < 3 * Hidden ? b % Hidden : b - 3 * Hidden)] =
            bias_acc;
    }
}

// grad = (sum of the workgroups' partials) / (sum of the batch weights).  A
// workgroup covers 64 parameters: wave w sums the partials of workgroups
// w, w + 8, w + 16, ... in order, then the eight sums are added in wave
// order (the weights alike), so gradients are deterministic.
constexpr int Reduce_waves = 8;

__global__ void __launch_bounds__(Reduce_waves * 64)
    reduce_partials(float const* partial, float const* wsum, uint32_t blocks, float* grad) {
    __shared__ float sum_s[Reduce_waves][64];
    __shared__ float sum_t[Reduce_waves];
    uint32_t const lane = threadIdx.x & 63;
    uint32_t const wave = threadIdx.x >> 6;
    uint32_t const p    = blockIdx.x * 64 + lane;
    float          s    = 0.f;
    float          t    = 0.f;
    if (p < uint32_t(Num_params)) {
        for (uint32_t b = wave; b < blocks; b += Reduce_waves) {
            s += partial[size_t(b) * Num_params + p];
            t += wsum[b];
        }
    }
    sum_s[wave][lane] = s;
    if (0 == lane) {
        sum_t[wave] = t;
    }
    __syncthreads();
    if (0 == wave && p < uint32_t(Num_params)) {
        float ss = 0.f;
        float tt = 0.f;
#pragma unroll
        for (int w = 0; w < Reduce_waves; ++w) {
            ss += sum_s[w][lane];
            tt += sum_t[w];
        }
        grad[p] = ss / (tt > 0.f ? tt : 1.f);
    }
}

// W^T bf16 position of parameter p, or -1 for a bias
__device__ __host__ inline int transposed_index(int p) {
    if (p < Off_b1) {
        return T_w1 + (p % Hidden) * Feat_pad + p / Hidden;
    }
    if (p >= Off_w2 && p < Off_b2) {
        int const e = p - Off_w2;
        return T_w2 + (e % Hidden) * Hidden + e / Hidden;
    }
    if (p >= Off_w3 && p < Off_b3) {
        int const e = p - Off_w3;
        return T_w3 + (e % Hidden) * Hidden + e / Hidden;
    }
    if (p >= Off_w4 && p < Off_b4) {
        int const e = p - Off_w4;
        return T_w4 + (e % Num_out) * Hidden + e / Num_out;
    }
    return -1;
}

__global__ void refresh_bf16(float const* params, uint16_t* wt) {
    int const p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= Num_params) {
        return;
    }
    int const t = transposed_index(p);
    if (t >= 0) {
        wt[t] = to_bf16(params[p]);
    }
}

// Training window of one Adam step: window = {first record, count, step t}.
// Records [k batch mod count, + batch) of the first min(*records, cap); with
// batch 0 the count is split into `steps` slices.  Empty windows take no
// step.  trained (optional) sums the records used.
__global__ void set_window(uint32_t const* records, uint32_t cap, uint32_t fixed_n, uint32_t k,
                           uint32_t steps, uint32_t batch, uint32_t* window, uint32_t* step,
                           unsigned long long* trained) {
    uint32_t off = 0;
    uint32_t nb  = fixed_n;
    if (records) {
        uint32_t const count = min(*records, cap);
        uint32_t const b     = batch ? min(batch, count) : (count + steps - 1) / max(steps, 1u);
        off                  = count ? uint32_t((uint64_t(k) * b) % count) : 0u;
        nb                   = count ? min(b, count - off) : 0u;
    }
    window[0] = off;
    window[1] = nb;
    if (nb) {
        *step += 1;
        if (trained) {
            *trained += nb;
        }
    }
    window[2] = *step;
}

// Adam with the bias corrections of the window's step (no step on an empty window).
__global__ void adam_kernel(float* params, float* m, float* v, float const* grad, float lr,
                            float b1, float b2, uint32_t const* window, float eps, uint16_t* wt) {
    int const p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= Num_params || 0 == window[1]) {
        return;
    }
    double const step = double(window[2]);
    float const  c1   = float(1.0 - pow(double(b1), step));
    float const  c2   = float(1.0 - pow(double(b2), step));
    float const g  = grad[p];
    float const mm = b1 * m[p] + (1.f - b1) * g;
    float const vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p]           = mm;
    v[p]           = vv;
    float const w  = params[p] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    params[p]      = w;
    int const t    = transposed_index(p);
    if (t >= 0) {
        wt[t] = to_bf16(w);
    }
}

uint32_t blocks_for(uint64_t n, uint32_t b) {
    uint64_t const g = (n + b - 1) / b;
    return uint32_t(g > 0 ? g : 1);
}

struct Pcg {
    uint64_t state, inc;
    uint32_t next() {
        uint64_t const old = state;
        state              = old * 6364136223846793005ull + (inc | 1);
        uint32_t const xrs = uint32_t(((old >> 18u) ^ old) >> 27u);
        uint32_t const rot = uint32_t(old >> 59u);
        return (xrs >> rot) | (xrs << ((0u - rot) & 31u));
    }
    double uniform() { return double(next()) * (1.0 / 4294967296.0); }
};

hipError_t refresh(GuideNet& net, hipStream_t stream) {
    hipLaunchKernelGGL(refresh_bf16, dim3(blocks_for(Num_params, 256)), dim3(256), 0, stream,
                       net.params, net.wt);
    return hipGetLastError();
}

}  // namespace

void GuideNet::release() {
    (void)hipFree(params);
    (void)hipFree(adam_m);
    (void)hipFree(adam_v);
    (void)hipFree(grad);
    (void)hipFree(partial);
    (void)hipFree(wsum);
    (void)hipFree(freq);
    (void)hipFree(wt);
    (void)hipFree(window);
    *this = GuideNet();
}

int32_t guide_create(uint64_t seed, uint32_t num_cus, GuideNet& out) {
    GuideNet n;
    n.partial_blocks = num_cus ? num_cus : 256;
    hipError_t e     = hipMalloc(&n.params, Num_params * sizeof(float));
    if (!e) e = hipMalloc(&n.adam_m, Num_params * sizeof(float));
    if (!e) e = hipMalloc(&n.adam_v, Num_params * sizeof(float));
    if (!e) e = hipMalloc(&n.grad, Num_params * sizeof(float));
    if (!e) e = hipMalloc(&n.partial, size_t(n.partial_blocks) * Num_params * sizeof(float));
    if (!e) e = hipMalloc(&n.wsum, size_t(n.partial_blocks) * sizeof(float));
    if (!e) e = hipMalloc(&n.freq, Num_in * Num_freq * sizeof(float));
    if (!e) {
        float f[Num_in * Num_freq];
        guide_default_frequencies(f);
        e = hipMemcpy(n.freq, f, sizeof(f), hipMemcpyHostToDevice);
    }
    if (!e) e = hipMalloc(&n.wt, Num_t * sizeof(uint16_t));
    if (!e) e = hipMalloc(&n.window, 8 * sizeof(uint32_t));
    if (e) {
        n.release();
        return fail(MI_ERR_HIP, "guide: %s", hipGetErrorString(e));
    }
    n.step_dev     = n.window + 4;
    n.scratch_step = n.window + 5;
    // He-uniform weights (fan-in of the unpadded features), zero biases
    std::vector<float> p(Num_params, 0.f);
    Pcg                rng{0, (seed << 1) | 1};
    rng.next();
    rng.state += seed;
    rng.next();
    struct L {
        int off, in, out, fan;
    } const layers[4] = {{Off_w1, Feat_pad, Hidden, Num_feat},
                         {Off_w2, Hidden, Hidden, Hidden},
                         {Off_w3, Hidden, Hidden, Hidden},
                         {Off_w4, Hidden, Num_out, Hidden}};
    for (L const& l : layers) {
        double const bound = std::sqrt(6.0 / l.fan);
        for (int i = 0; i < l.in; ++i) {
            for (int j = 0; j < l.out; ++j) {
                float const w = float((2.0 * rng.uniform() - 1.0) * bound);
                p[l.off + i * l.out + j] = (l.off == Off_w1 && i >= Num_feat) ? 0.f : w;
            }
        }
    }
    out = n;
    return guide_set_params(out, p.data());
}

void guide_default_frequencies(float* out) {
    for (int i = 0; i < Num_in; ++i) {
        double const step = i < 3 ? 0.5 : 0.25;  // position: half octaves, directions: quarter
        for (int j = 0; j < Num_freq; ++j) {
            out[i * Num_freq + j] = float(std::exp2(step * j));
        }
    }
}

int32_t guide_set_frequencies(GuideNet& net, float const* host_freq) {
    for (int k = 0; k < Num_in * Num_freq; ++k) {
        if (!std::isfinite(host_freq[k]) || std::fabs(host_freq[k]) > 4096.f) {
            return fail(MI_ERR_ARGS, "guide frequencies: entry %d is %g", k, double(host_freq[k]));
        }
    }
    hipError_t const e = hipMemcpy(net.freq, host_freq, Num_in * Num_freq * sizeof(float),
                                   hipMemcpyHostToDevice);
    if (e) {
        return fail(MI_ERR_HIP, "guide frequencies: %s", hipGetErrorString(e));
    }
    return MI_OK;
}

int32_t guide_set_params(GuideNet& net, float const* host_params) {
    hipError_t e = hipMemcpy(net.params, host_params, Num_params * sizeof(float), hipMemcpyHostToDevice);
    if (!e) e = hipMemset(net.adam_m, 0, Num_params * sizeof(float));
    if (!e) e = hipMemset(net.adam_v, 0, Num_params * sizeof(float));
    if (!e) e = hipMemset(net.window, 0, 8 * sizeof(uint32_t));  // step 0
    if (!e) e = refresh(net, nullptr);
    if (!e) e = hipDeviceSynchronize();
    if (e) {
        return fail(MI_ERR_HIP, "guide parameters: %s", hipGetErrorString(e));
    }
    return MI_OK;
}

int32_t guide_get_params(GuideNet const& net, float* host_params) {
    if (hipError_t const e = hipMemcpy(host_params, net.params, Num_params * sizeof(float),
                                       hipMemcpyDeviceToHost);
        e) {
        return fail(MI_ERR_HIP, "guide parameters: %s", hipGetErrorString(e));
    }
    return MI_OK;
}

int32_t guide_forward(GuideNet const& net, float const* x, uint32_t n, float* out,
                      hipStream_t stream) {
    if (0 == n) {
        return MI_OK;
    }
    // one workgroup per CU fits (LDS): more would reload the weights per round
    uint32_t const tiles  = blocks_for(n, 16);
    uint32_t const blocks = std::min<uint32_t>(blocks_for(tiles, Fwd_waves), net.partial_blocks);
    hipLaunchKernelGGL(forward_kernel, dim3(blocks), dim3(Fwd_waves * 64), 0, stream, net.wt,
                       net.params, net.freq, x, n, static_cast<uint32_t const*>(nullptr), out);
    if (hipError_t const e = hipGetLastError(); e) {
        return fail(MI_ERR_HIP, "guide forward: %s", hipGetErrorString(e));
    }
    return MI_OK;
}

int32_t guide_forward_device_count(GuideNet const& net, float const* x, uint32_t const* n,
                                   uint32_t max_n, float* out, hipStream_t stream) {
    if (0 == max_n) {
        return MI_OK;
    }
    uint32_t const tiles  = blocks_for(max_n, 16);
    uint32_t const blocks = std::min<uint32_t>(blocks_for(tiles, Fwd_waves), net.partial_blocks);
    hipLaunchKernelGGL(forward_kernel, dim3(blocks), dim3(Fwd_waves * 64), 0, stream, net.wt,
                       net.params, net.freq, x, max_n, n, out);
    if (hipError_t const e = hipGetLastError(); e) {
        return fail(MI_ERR_HIP, "guide forward: %s", hipGetErrorString(e));
    }
    return MI_OK;
}

namespace {

// gradient of the batch described by net.window (max_n bounds its count)
void enqueue_gradients(GuideNet& net, float const* x, float const* dirs, float const* weights,
                       uint32_t max_n, hipStream_t stream) {
    uint32_t const blocks = std::min(blocks_for(max_n, Chunk), net.partial_blocks);
    hipLaunchKernelGGL(gradient_kernel, dim3(blocks), dim3(Grad_waves * 64), 0, stream, net.wt,
                       net.params, net.freq, x, dirs, weights, 0u,
                       static_cast<uint32_t const*>(net.window), net.partial, net.wsum);
    hipLaunchKernelGGL(reduce_partials, dim3(blocks_for(Num_params, 64)), dim3(Reduce_waves * 64),
                       0, stream, net.partial, net.wsum, blocks, net.grad);
}

void enqueue_adam(GuideNet& net, float lr, hipStream_t stream) {
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(Num_params, 256)), dim3(256), 0, stream,
                       net.params, net.adam_m, net.adam_v, net.grad, lr, 0.9f, 0.99f,
                       static_cast<uint32_t const*>(net.window), 1e-8f, net.wt);
}

}  // namespace

int32_t guide_gradients(GuideNet& net, float const* x, float const* dirs, float const* weights,
                        uint32_t n, hipStream_t stream) {
    if (0 == n) {
        return fail(MI_ERR_ARGS, "guide gradients: empty batch");
    }
    // a fixed window of n records that takes no Adam step
    hipLaunchKernelGGL(set_window, dim3(1), dim3(1), 0, stream, static_cast<uint32_t const*>(nullptr),
                       0u, n, 0u, 1u, 0u, net.window, net.scratch_step,
                       static_cast<unsigned long long*>(nullptr));
    enqueue_gradients(net, x, dirs, weights, n, stream);
    if (hipError_t const e = hipGetLastError(); e) {
        return fail(MI_ERR_HIP, "guide gradients: %s", hipGetErrorString(e));
    }
    return MI_OK;
}

int32_t guide_adam_step(GuideNet& net, float const* x, float const* dirs, float const* weights,
                        uint32_t n, float lr, hipStream_t stream) {
    if (0 == n) {
        return fail(MI_ERR_ARGS, "guide adam: empty batch");
    }
    hipLaunchKernelGGL(set_window, dim3(1), dim3(1), 0, stream, static_cast<uint32_t const*>(nullptr),
                       0u, n, 0u, 1u, 0u, net.window, net.step_dev,
                       static_cast<unsigned long long*>(nullptr));
    enqueue_gradients(net, x, dirs, weights, n, stream);
    enqueue_adam(net, lr, stream);
    if (hipError_t const e = hipGetLastError(); e) {
        return fail(MI_ERR_HIP, "guide adam: %s", hipGetErrorString(e));
    }
    return MI_OK;
}

int32_t guide_train_records(GuideNet& net, float const* x, float const* dirs, float const* weights,
                            uint32_t const* num_records, uint32_t capacity, uint32_t steps,
                            uint32_t batch, float lr, unsigned long long* trained,
                            hipStream_t stream) {
    for (uint32_t k = 0; k < steps; ++k) {
        hipLaunchKernelGGL(set_window, dim3(1), dim3(1), 0, stream, num_records, capacity, 0u, k,
                           steps, batch, net.window, net.step_dev, trained);
        enqueue_gradients(net, x, dirs, weights, batch ? std::min(batch, capacity) : capacity, stream);
        enqueue_adam(net, lr, stream);
    }
    if (hipError_t const e = hipGetLastError(); e) {
        return fail(MI_ERR_HIP, "guide training: %s", hipGetErrorString(e));
    }
    return MI_OK;
}

}  // namespace mi
