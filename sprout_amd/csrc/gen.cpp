// Deterministic synthetic scenes and ray sets (host side).
//
// These produce the inputs of the BVH micro-benchmark: a displaced
// icosphere plus small random triangles, pinhole camera rays and cosine
// bounce rays from the camera hits.  Arithmetic is done in double and
// rounded to float once, so the same seed gives the same bits on any
// IEEE host.

#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

// PCG32 (XSH-RR), the generator sprout uses (base/random/generator.inl:16-49).
struct Pcg32 {
    uint64_t state;
    uint64_t inc;

    Pcg32(uint64_t seed, uint64_t stream) : state(0), inc((stream << 1u) | 1u) {
        next();
        state += seed;
        next();
    }

    uint32_t next() {
        uint64_t const old = state;
        state              = old * 6364136223846793005ull + inc;
        uint32_t const xorshifted = uint32_t(((old >> 18u) ^ old) >> 27u);
        uint32_t const rot        = uint32_t(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((~rot + 1u) & 31));
    }

    // [0, 1) with 32 random bits
    double uniform() { return double(next()) * (1.0 / 4294967296.0); }
};

uint32_t hash3(int32_t x, int32_t y, int32_t z, uint64_t seed) {
    uint64_t h = seed * 0x9E3779B97F4A7C15ull;
    h ^= uint64_t(uint32_t(x)) * 0xBF58476D1CE4E5B9ull;
    h = (h ^ (h >> 31)) * 0x94D049BB133111EBull;
    h ^= uint64_t(uint32_t(y)) * 0xD6E8FEB86659FD93ull;
    h = (h ^ (h >> 29)) * 0xBF58476D1CE4E5B9ull;
    h ^= uint64_t(uint32_t(z)) * 0x9E3779B97F4A7C15ull;
    h = (h ^ (h >> 32)) * 0x94D049BB133111EBull;
    return uint32_t(h >> 32);
}

// lattice value in [-1, 1]
double lattice(int32_t x, int32_t y, int32_t z, uint64_t seed) {
    return double(hash3(x, y, z, seed) >> 8) * (2.0 / 16777216.0) - 1.0;
}

double smooth(double t) {
    return t * t * (3.0 - 2.0 * t);
}

double value_noise(double x, double y, double z, uint64_t seed) {
    double const fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int32_t const ix = int32_t(fx), iy = int32_t(fy), iz = int32_t(fz);
    double const  tx = smooth(x - fx), ty = smooth(y - fy), tz = smooth(z - fz);

    double c[2][2][2];
    for (int k = 0; k < 2; ++k) {
        for (int j = 0; j < 2; ++j) {
            for (int i = 0; i < 2; ++i) {
                c[k][j][i] = lattice(ix + i, iy + j, iz + k, seed);
            }
        }
    }

    double r[2];
    for (int k = 0; k < 2; ++k) {
        double const a = c[k][0][0] + tx * (c[k][0][1] - c[k][0][0]);
        double const b = c[k][1][0] + tx * (c[k][1][1] - c[k][1][0]);
        r[k]           = a + ty * (b - a);
    }
    return r[0] + tz * (r[1] - r[0]);
}

// octaves of value noise, base frequency 2, gain 0.5, normalised to [-1, 1]
double fbm(double x, double y, double z, uint32_t octaves, uint64_t seed) {
    double sum = 0.0, norm = 0.0, amp = 1.0, freq = 2.0;
    for (uint32_t o = 0; o < octaves; ++o) {
        sum += amp * value_noise(x * freq, y * freq, z * freq, seed + o);
        norm += amp;
        amp *= 0.5;
        freq *= 2.0;
    }
    return norm > 0.0 ? sum / norm : 0.0;
}

struct D3 {
    double x, y, z;
};

D3 normalize(D3 v) {
    double const l = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    return {v.x / l, v.y / l, v.z / l};
}

void icosphere(uint32_t subdivisions, std::vector<D3>& verts, std::vector<uint32_t>& tris) {
    double const t = (1.0 + std::sqrt(5.0)) / 2.0;

    D3 const base[12] = {{-1, t, 0}, {1, t, 0},  {-1, -t, 0}, {1, -t, 0}, {0, -1, t},  {0, 1, t},
                         {0, -1, -t}, {0, 1, -t}, {t, 0, -1},  {t, 0, 1},  {-t, 0, -1}, {-t, 0, 1}};

    uint32_t const faces[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                                   {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                   {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                                   {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};

    verts.clear();
    for (auto const& v : base) {
        verts.push_back(normalize(v));
    }

    tris.clear();
    for (auto const& f : faces) {
        tris.insert(tris.end(), {f[0], f[1], f[2]});
    }

    for (uint32_t s = 0; s < subdivisions; ++s) {
        std::unordered_map<uint64_t, uint32_t> mids;
        mids.reserve(tris.size());

        auto midpoint = [&](uint32_t a, uint32_t b) -> uint32_t {
            uint64_t const key = a < b ? (uint64_t(a) << 32) | b : (uint64_t(b) << 32) | a;
            auto const     it  = mids.find(key);
            if (it != mids.end()) {
                return it->second;
            }
            D3 const& va = verts[a];
            D3 const& vb = verts[b];
            verts.push_back(normalize({0.5 * (va.x + vb.x), 0.5 * (va.y + vb.y), 0.5 * (va.z + vb.z)}));
            uint32_t const id = uint32_t(verts.size() - 1);
            mids.emplace(key, id);
            return id;
        };

        std::vector<uint32_t> next;
        next.reserve(tris.size() * 4);
        for (size_t i = 0; i < tris.size(); i += 3) {
            uint32_t const a = tris[i], b = tris[i + 1], c = tris[i + 2];
            uint32_t const ab = midpoint(a, b);
            uint32_t const bc = midpoint(b, c);
            uint32_t const ca = midpoint(c, a);
            next.insert(next.end(), {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca});
        }
        tris.swap(next);
    }
}

D3 random_direction(Pcg32& rng) {
    double const z   = 2.0 * rng.uniform() - 1.0;
    double const phi = 2.0 * M_PI * rng.uniform();
    double const r   = std::sqrt(std::max(0.0, 1.0 - z * z));
    return {r * std::cos(phi), r * std::sin(phi), z};
}

double gaussian(Pcg32& rng) {
    double const u1 = 1.0 - rng.uniform();  // (0, 1]
    double const u2 = rng.uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

float bits_offset(float p, int32_t offset) {
    int32_t i;
    std::memcpy(&i, &p, 4);
    i += p < 0.f ? -offset : offset;
    float r;
    std::memcpy(&r, &i, 4);
    return r;
}

// offset_ray, core/scene/ray_offset.inl:18-29
void offset_ray(float const p[3], float const n[3], float out[3]) {
    float const origin      = 1.f / 32.f;
    float const float_scale = 1.f / 65536.f;
    float const int_scale   = 256.f;
    for (int i = 0; i < 3; ++i) {
        int32_t const of_i = int32_t(int_scale * n[i]);
        out[i] = std::fabs(p[i]) < origin ? std::fmaf(float_scale, n[i], p[i]) : bits_offset(p[i], of_i);
    }
}

}  // namespace

extern "C" int32_t mi_gen_displaced_icosphere(uint32_t num_subdivisions, uint32_t octaves,
                                              float amplitude, uint64_t noise_seed,
                                              uint32_t num_random_triangles, float log_mu,
                                              float log_sigma, uint64_t random_seed,
                                              float* positions, uint32_t* num_vertices,
                                              uint32_t* indices, uint32_t* num_triangles) {
    if (num_subdivisions > 10 || !num_vertices || !num_triangles) {
        return mi::fail(MI_ERR_ARGS, "mi_gen_displaced_icosphere: bad arguments");
    }

    uint64_t const sphere_tris  = 20ull << (2 * num_subdivisions);
    uint64_t const sphere_verts = 10ull * (1ull << (2 * num_subdivisions)) + 2;
    uint64_t const total_tris   = sphere_tris + num_random_triangles;
    uint64_t const total_verts  = sphere_verts + 3ull * num_random_triangles;

    if (total_tris > 0xFFFFFFFFull / 3 || total_verts > 0xFFFFFFFFull / 3) {
        return mi::fail(MI_ERR_ARGS, "mi_gen_displaced_icosphere: mesh too large");
    }

    if (!positions || !indices) {
        *num_vertices  = uint32_t(total_verts);
        *num_triangles = uint32_t(total_tris);
        return MI_OK;
    }

    if (*num_vertices < total_verts || *num_triangles < total_tris) {
        return mi::fail(MI_ERR_ARGS, "mi_gen_displaced_icosphere: output buffers too small");
    }

    std::vector<D3>       verts;
    std::vector<uint32_t> tris;
    icosphere(num_subdivisions, verts, tris);

    for (size_t i = 0; i < verts.size(); ++i) {
        D3 const     v = verts[i];
        double const r = 1.0 + double(amplitude) * fbm(v.x, v.y, v.z, octaves, noise_seed);
        positions[3 * i + 0] = float(v.x * r);
        positions[3 * i + 1] = float(v.y * r);
        positions[3 * i + 2] = float(v.z * r);
    }
    std::memcpy(indices, tris.data(), tris.size() * sizeof(uint32_t));

    Pcg32 rng(random_seed, 0x5eed);

    uint32_t v = uint32_t(verts.size());
    for (uint32_t t = 0; t < num_random_triangles; ++t) {
        D3 const c = {rng.uniform(), rng.uniform(), rng.uniform()};
        for (uint32_t k = 0; k < 3; ++k) {
            D3 const     d   = random_direction(rng);
            double const len = std::exp(double(log_mu) + double(log_sigma) * gaussian(rng));
            positions[3 * (v + k) + 0] = float(c.x + len * d.x);
            positions[3 * (v + k) + 1] = float(c.y + len * d.y);
            positions[3 * (v + k) + 2] = float(c.z + len * d.z);
        }
        size_t const o = tris.size() + 3 * size_t(t);
        indices[o + 0] = v;
        indices[o + 1] = v + 1;
        indices[o + 2] = v + 2;
        v += 3;
    }

    *num_vertices  = uint32_t(total_verts);
    *num_triangles = uint32_t(total_tris);
    return MI_OK;
}

extern "C" int32_t mi_gen_camera_rays(uint32_t width, uint32_t height, float const eye[3],
                                      float const target[3], float const up[3], float fov_deg,
                                      mi_ray* rays) {
    if (!width || !height || !eye || !target || !up || !rays || !(fov_deg > 0.f) ||
        !(fov_deg < 180.f)) {
        return mi::fail(MI_ERR_ARGS, "mi_gen_camera_rays: bad arguments");
    }

    D3 const e   = {eye[0], eye[1], eye[2]};
    D3 const fwd = normalize({target[0] - e.x, target[1] - e.y, target[2] - e.z});
    D3 const u0  = {up[0], up[1], up[2]};
    // right = fwd x up, true up = right x fwd
    D3 const right = normalize({fwd.y * u0.z - fwd.z * u0.y, fwd.z * u0.x - fwd.x * u0.z,
                                fwd.x * u0.y - fwd.y * u0.x});
    D3 const tup = {right.y * fwd.z - right.z * fwd.y, right.z * fwd.x - right.x * fwd.z,
                    right.x * fwd.y - right.y * fwd.x};

    double const th     = std::tan(0.5 * double(fov_deg) * M_PI / 180.0);
    double const aspect = double(width) / double(height);

    for (uint32_t y = 0; y < height; ++y) {
        for (uint32_t x = 0; x < width; ++x) {
            double const sx = (2.0 * (x + 0.5) / width - 1.0) * th * aspect;
            double const sy = (1.0 - 2.0 * (y + 0.5) / height) * th;
            D3 const     d  = normalize({fwd.x + sx * right.x + sy * tup.x,
                                         fwd.y + sx * right.y + sy * tup.y,
                                         fwd.z + sx * right.z + sy * tup.z});
            mi_ray& r      = rays[size_t(y) * width + x];
            r.origin[0]    = eye[0];
            r.origin[1]    = eye[1];
            r.origin[2]    = eye[2];
            r.min_t        = 0.f;
            r.direction[0] = float(d.x);
            r.direction[1] = float(d.y);
            r.direction[2] = float(d.z);
            r.max_t        = 3.40282347e+38f;
        }
    }
    return MI_OK;
}

extern "C" int32_t mi_gen_bounce_rays(mi_ray const* in_rays, mi_hit const* hits, uint32_t num_in,
                                      float const* positions, uint32_t position_stride,
                                      uint32_t const* indices, uint64_t seed, mi_ray* out_rays,
                                      uint32_t num_out) {
    if (!in_rays || !hits || !positions || !indices || !out_rays || position_stride < 3) {
        return mi::fail(MI_ERR_ARGS, "mi_gen_bounce_rays: bad arguments");
    }

    std::vector<uint32_t> hit_ids;
    hit_ids.reserve(num_in);
    for (uint32_t i = 0; i < num_in; ++i) {
        if (hits[i].triangle != MI_NULL) {
            hit_ids.push_back(i);
        }
    }
    if (hit_ids.empty()) {
        return mi::fail(MI_ERR_ARGS, "mi_gen_bounce_rays: no input ray hits");
    }

    for (uint32_t i = 0; i < num_out; ++i) {
        uint32_t const src = hit_ids[i % hit_ids.size()];
        mi_ray const&  ir  = in_rays[src];
        mi_hit const&  h   = hits[src];

        float const* pa = positions + size_t(indices[3 * size_t(h.triangle) + 0]) * position_stride;
        float const* pb = positions + size_t(indices[3 * size_t(h.triangle) + 1]) * position_stride;
        float const* pc = positions + size_t(indices[3 * size_t(h.triangle) + 2]) * position_stride;

        double const w = 1.0 - double(h.u) - double(h.v);
        float        p[3];
        for (int k = 0; k < 3; ++k) {
            p[k] = float(w * pa[k] + double(h.u) * pb[k] + double(h.v) * pc[k]);
        }

        D3 const e1 = {double(pb[0]) - pa[0], double(pb[1]) - pa[1], double(pb[2]) - pa[2]};
        D3 const e2 = {double(pc[0]) - pa[0], double(pc[1]) - pa[1], double(pc[2]) - pa[2]};
        D3       n  = normalize({e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z,
                                 e1.x * e2.y - e1.y * e2.x});
        if (n.x * ir.direction[0] + n.y * ir.direction[1] + n.z * ir.direction[2] > 0.0) {
            n = {-n.x, -n.y, -n.z};
        }

        float const nf[3] = {float(n.x), float(n.y), float(n.z)};
        float       o[3];
        offset_ray(p, nf, o);

        // orthonormal basis around n (Duff et al. 2017)
        double const s  = std::copysign(1.0, n.z);
        double const a  = -1.0 / (s + n.z);
        double const b  = n.x * n.y * a;
        D3 const     t0 = {1.0 + s * n.x * n.x * a, s * b, -s * n.x};
        D3 const     t1 = {b, s + n.y * n.y * a, -n.y};

        Pcg32        rng(seed, i);
        double const u1  = rng.uniform();
        double const u2  = rng.uniform();
        double const r   = std::sqrt(u1);
        double const phi = 2.0 * M_PI * u2;
        double const lx = r * std::cos(phi), ly = r * std::sin(phi);
        double const lz = std::sqrt(std::max(0.0, 1.0 - u1));
        D3 const     d  = normalize({lx * t0.x + ly * t1.x + lz * n.x, lx * t0.y + ly * t1.y + lz * n.y,
                                     lx * t0.z + ly * t1.z + lz * n.z});

        mi_ray& out      = out_rays[i];
        out.origin[0]    = o[0];
        out.origin[1]    = o[1];
        out.origin[2]    = o[2];
        out.min_t        = 0.f;
        out.direction[0] = float(d.x);
        out.direction[1] = float(d.y);
        out.direction[2] = float(d.z);
        out.max_t        = 3.40282347e+38f;
    }
    return MI_OK;
}

namespace {

struct Room {
    std::vector<float>    positions;
    std::vector<uint32_t> indices;
    std::vector<uint32_t> materials;

    // Axis-aligned quad with corners p0..p3 (in order around it), wound so
    // that cross(b - a, c - a) points along n.
    void quad(D3 p0, D3 p1, D3 p2, D3 p3, D3 n, uint32_t material) {
        D3 const e1 = {p1.x - p0.x, p1.y - p0.y, p1.z - p0.z};
        D3 const e2 = {p2.x - p0.x, p2.y - p0.y, p2.z - p0.z};
        D3 const c  = {e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z,
                       e1.x * e2.y - e1.y * e2.x};
        bool const flip = c.x * n.x + c.y * n.y + c.z * n.z < 0.0;
        uint32_t const base = uint32_t(positions.size() / 3);
        for (D3 const& p : {p0, p1, p2, p3}) {
            positions.push_back(float(p.x));
            positions.push_back(float(p.y));
            positions.push_back(float(p.z));
        }
        uint32_t const t[6] = {0, 1, 2, 0, 2, 3};
        uint32_t const f[6] = {0, 2, 1, 0, 3, 2};
        for (int k = 0; k < 6; ++k) {
            indices.push_back(base + (flip ? f[k] : t[k]));
        }
        materials.push_back(material);
        materials.push_back(material);
    }

    // Box [lo, hi] with outward faces.
    void box(D3 lo, D3 hi, uint32_t material) {
        quad({lo.x, lo.y, lo.z}, {hi.x, lo.y, lo.z}, {hi.x, lo.y, hi.z}, {lo.x, lo.y, hi.z}, {0, -1, 0}, material);
        quad({lo.x, hi.y, lo.z}, {hi.x, hi.y, lo.z}, {hi.x, hi.y, hi.z}, {lo.x, hi.y, hi.z}, {0, 1, 0}, material);
        quad({lo.x, lo.y, lo.z}, {lo.x, hi.y, lo.z}, {lo.x, hi.y, hi.z}, {lo.x, lo.y, hi.z}, {-1, 0, 0}, material);
        quad({hi.x, lo.y, lo.z}, {hi.x, hi.y, lo.z}, {hi.x, hi.y, hi.z}, {hi.x, lo.y, hi.z}, {1, 0, 0}, material);
        quad({lo.x, lo.y, lo.z}, {hi.x, lo.y, lo.z}, {hi.x, hi.y, lo.z}, {lo.x, hi.y, lo.z}, {0, 0, -1}, material);
        quad({lo.x, lo.y, hi.z}, {hi.x, lo.y, hi.z}, {hi.x, hi.y, hi.z}, {lo.x, hi.y, hi.z}, {0, 0, 1}, material);
    }
};

Room make_room(float slit) {
    Room r;
    // floor, ceiling, back wall (grey), left (red), right (green); +z open
    r.quad({-1, 0, -1}, {1, 0, -1}, {1, 0, 1}, {-1, 0, 1}, {0, 1, 0}, 0);
    r.quad({-1, 2, -1}, {1, 2, -1}, {1, 2, 1}, {-1, 2, 1}, {0, -1, 0}, 0);
    r.quad({-1, 0, -1}, {1, 0, -1}, {1, 2, -1}, {-1, 2, -1}, {0, 0, 1}, 0);
    r.quad({-1, 0, -1}, {-1, 2, -1}, {-1, 2, 1}, {-1, 0, 1}, {1, 0, 0}, 1);
    r.quad({1, 0, -1}, {1, 2, -1}, {1, 2, 1}, {1, 0, 1}, {-1, 0, 0}, 2);
    // a tall and a short box
    r.box({-0.65, 0.0, -0.65}, {-0.15, 1.2, -0.15}, 0);
    r.box({0.15, 0.0, -0.2}, {0.65, 0.6, 0.3}, 0);
    // ceiling light, facing down
    r.quad({-0.25, 1.98, -0.25}, {0.25, 1.98, -0.25}, {0.25, 1.98, 0.25}, {-0.25, 1.98, 0.25},
           {0, -1, 0}, 3);
    if (slit > 0.f) {
        // enclosure below the light: bottom plate and side walls facing out,
        // the -z wall open for y in [1.7, 1.7 + slit]
        double const lo = 1.7;
        double const s  = 0.3;
        r.quad({-s, lo, -s}, {s, lo, -s}, {s, lo, s}, {-s, lo, s}, {0, -1, 0}, 0);
        r.quad({-s, lo, -s}, {-s, 2, -s}, {-s, 2, s}, {-s, lo, s}, {-1, 0, 0}, 0);
        r.quad({s, lo, -s}, {s, 2, -s}, {s, 2, s}, {s, lo, s}, {1, 0, 0}, 0);
        r.quad({-s, lo, s}, {s, lo, s}, {s, 2, s}, {-s, 2, s}, {0, 0, 1}, 0);
        double const top = lo + double(slit);
        r.quad({-s, top, -s}, {s, top, -s}, {s, 2, -s}, {-s, 2, -s}, {0, 0, -1}, 0);
    }
    return r;
}

}  // namespace

extern "C" int32_t mi_gen_room(float slit, float* positions, uint32_t* num_vertices,
                               uint32_t* indices, uint32_t* num_triangles,
                               uint32_t* triangle_materials, mi_material* materials,
                               uint32_t* num_materials) {
    if (!num_vertices || !num_triangles || !num_materials || !(slit >= 0.f) || !(slit < 0.3f)) {
        return mi::fail(MI_ERR_ARGS, "mi_gen_room: bad arguments");
    }
    Room const r = make_room(slit);
    *num_vertices  = uint32_t(r.positions.size() / 3);
    *num_triangles = uint32_t(r.indices.size() / 3);
    *num_materials = 4;
    if (positions) {
        std::memcpy(positions, r.positions.data(), r.positions.size() * sizeof(float));
    }
    if (indices) {
        std::memcpy(indices, r.indices.data(), r.indices.size() * sizeof(uint32_t));
    }
    if (triangle_materials) {
        std::memcpy(triangle_materials, r.materials.data(), r.materials.size() * sizeof(uint32_t));
    }
    if (materials) {
        mi_material const grey  = {{0.75f, 0.75f, 0.75f}, {0, 0, 0}, 1.f, 0.f, 1.46f, MI_MATERIAL_LAMBERT};
        mi_material const red   = {{0.63f, 0.06f, 0.05f}, {0, 0, 0}, 1.f, 0.f, 1.46f, MI_MATERIAL_LAMBERT};
        mi_material const green = {{0.14f, 0.45f, 0.09f}, {0, 0, 0}, 1.f, 0.f, 1.46f, MI_MATERIAL_LAMBERT};
        mi_material const light = {{0, 0, 0}, {17.f, 12.f, 4.f}, 1.f, 0.f, 1.f, MI_MATERIAL_LIGHT};
        materials[0] = grey;
        materials[1] = red;
        materials[2] = green;
        materials[3] = light;
    }
    return MI_OK;
}

namespace {

Room make_cloud_scene() {
    Room r;
    // ground, grey, facing up
    r.quad({-4, 0, -4}, {4, 0, -4}, {4, 0, 4}, {-4, 0, 4}, {0, 1, 0}, 0);
    // emissive quad above the cloud, facing down: 25 x 20 cells, 1000 triangles
    double const x0 = -3.0, x1 = 3.0, z0 = -3.0, z1 = 3.0, y = 3.5;
    uint32_t const nx = 25, nz = 20;
    for (uint32_t j = 0; j < nz; ++j) {
        for (uint32_t i = 0; i < nx; ++i) {
            double const ax = x0 + (x1 - x0) * double(i) / nx;
            double const bx = x0 + (x1 - x0) * double(i + 1) / nx;
            double const az = z0 + (z1 - z0) * double(j) / nz;
            double const bz = z0 + (z1 - z0) * double(j + 1) / nz;
            r.quad({ax, y, az}, {bx, y, az}, {bx, y, bz}, {ax, y, bz}, {0, -1, 0}, 1);
        }
    }
    return r;
}

}  // namespace

extern "C" int32_t mi_gen_cloud_scene(float* positions, uint32_t* num_vertices, uint32_t* indices,
                                      uint32_t* num_triangles, uint32_t* triangle_materials,
                                      mi_material* materials, uint32_t* num_materials) {
    if (!num_vertices || !num_triangles || !num_materials) {
        return mi::fail(MI_ERR_ARGS, "mi_gen_cloud_scene: bad arguments");
    }
    Room const r = make_cloud_scene();
    *num_vertices  = uint32_t(r.positions.size() / 3);
    *num_triangles = uint32_t(r.indices.size() / 3);
    *num_materials = 2;
    if (positions) {
        std::memcpy(positions, r.positions.data(), r.positions.size() * sizeof(float));
    }
    if (indices) {
        std::memcpy(indices, r.indices.data(), r.indices.size() * sizeof(uint32_t));
    }
    if (triangle_materials) {
        std::memcpy(triangle_materials, r.materials.data(), r.materials.size() * sizeof(uint32_t));
    }
    if (materials) {
        mi_material const ground = {{0.5f, 0.5f, 0.5f}, {0, 0, 0}, 1.f, 0.f, 1.46f, MI_MATERIAL_LAMBERT};
        mi_material const sky    = {{0, 0, 0}, {4.f, 4.5f, 5.f}, 1.f, 0.f, 1.f, MI_MATERIAL_LIGHT};
        materials[0] = ground;
        materials[1] = sky;
    }
    return MI_OK;
}

extern "C" int32_t mi_gen_fbm_density(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t octaves,
                                      uint64_t seed, float max_density, float* density) {
    if (!density || 0 == nx || 0 == ny || 0 == nz || nx > 4096 || ny > 4096 || nz > 4096 ||
        0 == octaves || octaves > 16 || !(max_density >= 0.f)) {
        return mi::fail(MI_ERR_ARGS, "mi_gen_fbm_density: bad arguments");
    }
    auto slab = [&](uint32_t z_begin, uint32_t z_end) {
        for (uint32_t z = z_begin; z < z_end; ++z) {
            for (uint32_t y = 0; y < ny; ++y) {
                for (uint32_t x = 0; x < nx; ++x) {
                    // voxel centre in [0, 1]^3, offset to the cube centre
                    double const u = (double(x) + 0.5) / nx - 0.5;
                    double const v = (double(y) + 0.5) / ny - 0.5;
                    double const w = (double(z) + 0.5) / nz - 0.5;
                    double const n = fbm(u * 2.0 + 7.31, v * 2.0 + 3.17, w * 2.0 + 5.53, octaves, seed);
                    // spherical envelope keeps the cloud clear of the box faces
                    double const r2  = 4.0 * (u * u + v * v + w * w);
                    double const env = std::max(0.0, 1.0 - r2);
                    double const c   = std::min(1.0, std::max(0.0, (0.5 * n + 0.5) * 2.2 * env - 0.35) * 2.0);
                    density[(size_t(z) * ny + y) * nx + x] = float(double(max_density) * c);
                }
            }
        }
    };
    uint32_t const threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    uint32_t const per = (nz + threads - 1) / threads;
    for (uint32_t t = 0; t < threads; ++t) {
        uint32_t const b = std::min(nz, t * per);
        uint32_t const e = std::min(nz, b + per);
        if (b < e) {
            pool.emplace_back(slab, b, e);
        }
    }
    for (auto& th : pool) {
        th.join();
    }
    return MI_OK;
}
