// Host-side scene generators for the synthetic workloads.
//
// Random triangles are seeded SplitMix64 -> PCG32 streams. Terrain follows
// RAPTor/terrain_generator/height_map.h:16-43 (fBm: octaves of value noise,
// frequency doubling, amplitude times persistence) over the hash value noise
// of common/perlin_noise_2d.cc:3-47, triangulated two triangles per cell as
// RAPTor/terrain_generator/main.cc:163-175 lays them out.
#include <cmath>
#include <cstdint>

#include "../../include/raptor_rt.h"

namespace {

uint64_t sm64_next(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

class Pcg32 {
  public:
    Pcg32(uint64_t initstate, uint64_t initseq) {
        inc_ = (initseq << 1u) | 1u;
        step();
        state_ += initstate;
        step();
    }
    uint32_t step() {
        const uint64_t old = state_;
        state_ = old * 6364136223846793005ull + inc_;
        const uint32_t xs = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = static_cast<uint32_t>(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    // 24 random bits -> [0, 1)
    float uniform() { return static_cast<float>(step() >> 8) * (1.0f / 16777216.0f); }

  private:
    uint64_t state_ = 0, inc_ = 1;
};

// Two independent PCG32 streams derived from one seed: geometry and albedo.
struct Streams {
    explicit Streams(uint64_t seed) {
        uint64_t s = seed;
        const uint64_t a = sm64_next(&s);
        const uint64_t b = sm64_next(&s);
        const uint64_t c = sm64_next(&s);
        const uint64_t d = sm64_next(&s);
        geo = new Pcg32(a, b);
        alb = new Pcg32(c, d);
    }
    ~Streams() {
        delete geo;
        delete alb;
    }
    Pcg32 *geo, *alb;
};

// common/perlin_noise_2d.cc:3-9 — integer hash noise in [-1, 1]. Signed
// overflow of the reference's int arithmetic is taken as 32-bit wrap-around.
float noise(int x, int y, int seed) {
    uint32_t n = static_cast<uint32_t>(x) + static_cast<uint32_t>(y) * 57u + static_cast<uint32_t>(seed);
    n = (n << 13) ^ n;
    const uint32_t v = (n * (n * n * 15731u + 789221u) + 1376312589u) & 0x7fffffffu;
    return 1.0f - static_cast<float>(static_cast<int32_t>(v)) / 1073741824.0f;
}

// common/perlin_noise_2d.cc:12-18
float smooth_noise(int x, int y, int seed) {
    const float corners =
        (((noise(x - 1, y - 1, seed) + noise(x + 1, y - 1, seed)) + noise(x - 1, y + 1, seed)) + noise(x + 1, y + 1, seed)) /
        16.0f;
    const float sides =
        (((noise(x - 1, y, seed) + noise(x + 1, y, seed)) + noise(x, y - 1, seed)) + noise(x, y + 1, seed)) / 8.0f;
    const float center = noise(x, y, seed) / 4.0f;
    return (corners + sides) + center;
}

// common/perlin_noise_2d.cc:21-27 — cosine interpolation.
float interpolate(float a, float b, float x) {
    const float pi = 3.1415926535897932384626433f;  // common/common.cc:10
    const float ft = x * pi;
    const float f = (1.0f - std::cos(ft)) * 0.5f;
    return (a * (1.0f - f)) + (b * f);
}

// common/perlin_noise_2d.cc:30-47
float interpolated_noise(float x, float y, int seed) {
    const int wx = static_cast<int>(static_cast<float>(static_cast<int>(x)) + 0.5f);
    const float px = x - static_cast<float>(wx);
    const int wy = static_cast<int>(static_cast<float>(static_cast<int>(y)) + 0.5f);
    const float py = y - static_cast<float>(wy);
    const float v1 = smooth_noise(wx, wy, seed);
    const float v2 = smooth_noise(wx + 1, wy, seed);
    const float v3 = smooth_noise(wx, wy + 1, seed);
    const float v4 = smooth_noise(wx + 1, wy + 1, seed);
    const float i1 = interpolate(v1, v2, px);
    const float i2 = interpolate(v3, v4, px);
    return interpolate(i1, i2, py);
}

}  // namespace

extern "C" int64_t rt_gen_random_triangles(uint64_t seed, int64_t n, float spread, float albedo_lo, float albedo_hi,
                                           float *abc, float *albedo) {
    if (n < 0) return -1;
    if (!abc) return n;
    Streams st(seed);
    for (int64_t i = 0; i < n; ++i) {
        const float cx = st.geo->uniform(), cy = st.geo->uniform(), cz = st.geo->uniform();
        const float c[3] = {cx, cy, cz};
        for (int v = 0; v < 3; ++v) {
            for (int k = 0; k < 3; ++k) {
                const float off = ((st.geo->uniform() * 2.0f) - 1.0f) * spread;
                abc[9 * i + 3 * v + k] = c[k] + off;
            }
        }
        if (albedo) {
            for (int k = 0; k < 3; ++k) albedo[3 * i + k] = albedo_lo + ((albedo_hi - albedo_lo) * st.alb->uniform());
        }
    }
    return n;
}

extern "C" int64_t rt_gen_terrain(int32_t grid, float amplitude, float zoom, float persistence, int32_t octaves,
                                  int32_t noise_seed, uint64_t albedo_seed, float albedo_lo, float albedo_hi,
                                  float *abc, float *albedo) {
    if (grid < 2) return -1;
    const int64_t g = grid;
    const int64_t ntri = 2 * (g - 1) * (g - 1);
    if (!abc) return ntri;
    // height_map::generate(xs, ys) with xs = ys = 1 / (grid - 1).
    const float xs = 1.0f / static_cast<float>(grid - 1);
    const float ys = xs;
    float *verts = new float[3 * g * g];
    for (int64_t i = 0; i < g; ++i) {
        const float x = static_cast<float>(i) * zoom;
        for (int64_t j = 0; j < g; ++j) {
            const float y = static_cast<float>(j) * zoom;
            float frequency = 1.0f;
            float amp = amplitude;
            float total = 0.0f;
            for (int k = 0; k < octaves; ++k) {
                total += (interpolated_noise(x * frequency, y * frequency, noise_seed) * amp);
                frequency *= 2.0f;
                amp *= persistence;
            }
            float *v = verts + 3 * (i * g + j);
            v[0] = static_cast<float>(static_cast<int>(i)) * xs;
            v[1] = total;
            v[2] = static_cast<float>(static_cast<int>(j)) * ys;
        }
    }
    // Two triangles per cell: (r+j, r+j+1, r'+j+1) and (r'+j, r+j, r'+j+1).
    int64_t t = 0;
    for (int64_t i = 0; i < g - 1; ++i) {
        const int64_t r0 = i * g, r1 = r0 + g;
        for (int64_t j = 0; j < g - 1; ++j) {
            const int64_t ids[2][3] = {{r0 + j, r0 + j + 1, r1 + j + 1}, {r1 + j, r0 + j, r1 + j + 1}};
            for (int q = 0; q < 2; ++q) {
                for (int v = 0; v < 3; ++v) {
                    for (int k = 0; k < 3; ++k) abc[9 * t + 3 * v + k] = verts[3 * ids[q][v] + k];
                }
                ++t;
            }
        }
    }
    delete[] verts;
    if (albedo) {
        Streams st(albedo_seed);
        for (int64_t i = 0; i < ntri; ++i) {
            for (int k = 0; k < 3; ++k) albedo[3 * i + k] = albedo_lo + ((albedo_hi - albedo_lo) * st.alb->uniform());
        }
    }
    return ntri;
}
