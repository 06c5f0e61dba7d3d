// Oracle restatement of the noise FastNoise computes for SphereNoiseEditor (test
// infrastructure only): the seeded permutation (FastNoise.cpp:207-225, std::mt19937_64), the
// fractal bounding (227-237), the 3D Perlin gradient noise with quintic interpolation (195-205,
// 328-333, 841-888) and its FBM sum (760-797). Strict IEEE float in the source's order.
#include <cstdint>
#include <random>

#include "oracle_core.hpp"

namespace orc {

namespace {
// the 12 gradient directions (FastNoise.cpp:48-66)
const float kGradX[12] = {1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
const float kGradY[12] = {1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
const float kGradZ[12] = {0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};

int fast_floor(float f) { return f >= 0 ? (int)f : (int)f - 1; }
float lerp(float a, float b, float t) { return a + t * (b - a); }
float quintic(float t) { return t * t * t * (t * (t * 6 - 15) + 10); }
}  // namespace

void PerlinFbm::set_seed(int seed) {
    std::mt19937_64 gen((std::mt19937_64::result_type)(int64_t)seed);
    for (int i = 0; i < 256; ++i) perm[i] = (u8)i;
    for (int j = 0; j < 256; ++j) {
        const int k = (int)(gen() % (uint64_t)(256 - j)) + j;
        const u8 l = perm[j];
        perm[j] = perm[j + 256] = perm[k];
        perm[k] = l;
        perm12[j] = perm12[j + 256] = (u8)(perm[j] % 12);
    }
}

void PerlinFbm::set_octaves(int n) {
    octaves = n;
    float amp = gain, ampFractal = 1.0f;
    for (int i = 1; i < octaves; ++i) {
        ampFractal += amp;
        amp *= gain;
    }
    bounding = 1.0f / ampFractal;
}

float PerlinFbm::single(u8 offset, float x, float y, float z) const {
    const int x0 = fast_floor(x), y0 = fast_floor(y), z0 = fast_floor(z);
    const int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
    const float xs = quintic(x - (float)x0), ys = quintic(y - (float)y0), zs = quintic(z - (float)z0);
    const float xd0 = x - (float)x0, yd0 = y - (float)y0, zd0 = z - (float)z0;
    const float xd1 = xd0 - 1, yd1 = yd0 - 1, zd1 = zd0 - 1;
    auto grad = [&](int gx, int gy, int gz, float dx, float dy, float dz) {
        const u8 lut = perm12[(gx & 0xff) + perm[(gy & 0xff) + perm[(gz & 0xff) + offset]]];
        return dx * kGradX[lut] + dy * kGradY[lut] + dz * kGradZ[lut];
    };
    const float xf00 = lerp(grad(x0, y0, z0, xd0, yd0, zd0), grad(x1, y0, z0, xd1, yd0, zd0), xs);
    const float xf10 = lerp(grad(x0, y1, z0, xd0, yd1, zd0), grad(x1, y1, z0, xd1, yd1, zd0), xs);
    const float xf01 = lerp(grad(x0, y0, z1, xd0, yd0, zd1), grad(x1, y0, z1, xd1, yd0, zd1), xs);
    const float xf11 = lerp(grad(x0, y1, z1, xd0, yd1, zd1), grad(x1, y1, z1, xd1, yd1, zd1), xs);
    const float yf0 = lerp(xf00, xf10, ys), yf1 = lerp(xf01, xf11, ys);
    return lerp(yf0, yf1, zs);
}

float PerlinFbm::fractal(float x, float y, float z) const {
    x *= frequency;
    y *= frequency;
    z *= frequency;
    float sum = single(perm[0], x, y, z), amp = 1;
    for (int i = 1; i < octaves; ++i) {
        x *= lacunarity;
        y *= lacunarity;
        z *= lacunarity;
        amp *= gain;
        sum += single(perm[i], x, y, z) * amp;
    }
    return sum * bounding;
}

}  // namespace orc
