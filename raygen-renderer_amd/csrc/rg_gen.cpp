// rg_gen.cpp — workload generators (not part of the traced path).
//
// The displaced icosphere of the traversal microbenchmark: a subdivided
// icosahedron on the unit sphere, each vertex pushed along its radius by
// 1 + amp*(2*fbm(p) - 1).  fbm uses the value-noise basis of the reference's
// participating media (src/raygen/medium.cpp:125-185: integer-hash lattice,
// smoothstep trilinear interpolation, octave sum normalised by the amplitude
// sum) with 3 octaves, base frequency 4, gain 0.5, lacunarity 2, and a lattice
// offset drawn from the seed.  Shared edges are split with the same binary32
// operations from both sides, so the mesh is watertight.
#include <cmath>
#include <cstring>
#include <vector>

#include "rg_internal.h"

namespace {

struct V3 {
    float x, y, z;
};

inline float lattice(int x, int y, int z) {
    uint32_t h = (uint32_t)x * 374761393u + (uint32_t)y * 668265263u + (uint32_t)z * 2147483647u;
    h = (h ^ (h >> 13)) * 1274126177u;
    h ^= h >> 16;
    return (float)(h & 0x00FFFFFFu) / (float)0x01000000u;
}

inline float fade(float t) { return t * t * (3.0f - 2.0f * t); }

float value_noise(float x, float y, float z) {
    const int xi = (int)floorf(x), yi = (int)floorf(y), zi = (int)floorf(z);
    const float fx = fade(x - (float)xi), fy = fade(y - (float)yi), fz = fade(z - (float)zi);
    float c[2][2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int d = 0; d < 2; d++) c[a][b][d] = lattice(xi + a, yi + b, zi + d);
    const float x00 = c[0][0][0] * (1.0f - fx) + c[1][0][0] * fx;
    const float x10 = c[0][1][0] * (1.0f - fx) + c[1][1][0] * fx;
    const float x01 = c[0][0][1] * (1.0f - fx) + c[1][0][1] * fx;
    const float x11 = c[0][1][1] * (1.0f - fx) + c[1][1][1] * fx;
    const float y0 = x00 * (1.0f - fy) + x10 * fy;
    const float y1 = x01 * (1.0f - fy) + x11 * fy;
    return y0 * (1.0f - fz) + y1 * fz;
}

float fbm3(float x, float y, float z) {
    float f = 4.0f, amp = 1.0f, sum = 0.0f, norm = 0.0f;
    for (int o = 0; o < 3; o++) {
        sum += amp * value_noise(x * f, y * f, z * f);
        norm += amp;
        amp *= 0.5f;
        f *= 2.0f;
    }
    return sum / norm;
}

void philox_host(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

inline V3 unit(V3 p) {
    const float l = sqrtf(p.x * p.x + p.y * p.y + p.z * p.z);
    return {p.x / l, p.y / l, p.z / l};
}

inline V3 mid(V3 a, V3 b) { return unit({(a.x + b.x) * 0.5f, (a.y + b.y) * 0.5f, (a.z + b.z) * 0.5f}); }

}  // namespace

extern "C" int rg_gen_icosphere(uint32_t level, float amp, uint32_t seed, float *out_v9) {
    using rg::set_error;
    if (!out_v9) return set_error(RG_E_INVALID, "out_v9 is null");
    if (level > 10) return set_error(RG_E_INVALID, "level must be <= 10");
    const float t = (1.0f + sqrtf(5.0f)) * 0.5f;
    const V3 base[12] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                         {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    static const int faces[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                                     {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                     {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                                     {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
    std::vector<V3> cur, nxt;
    for (auto &f : faces)
        for (int k = 0; k < 3; k++) cur.push_back(unit(base[f[k]]));
    for (uint32_t l = 0; l < level; l++) {
        nxt.resize(cur.size() * 4);
        for (size_t i = 0; i < cur.size() / 3; i++) {
            const V3 a = cur[3 * i], b = cur[3 * i + 1], c = cur[3 * i + 2];
            const V3 ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
            const V3 tri[12] = {a, ab, ca, ab, b, bc, ca, bc, c, ab, bc, ca};
            std::memcpy(&nxt[12 * i], tri, sizeof tri);
        }
        cur.swap(nxt);
    }
    uint32_t ctr[4] = {0, 0, 0, 0};
    philox_host(ctr, seed, 0);
    const float off[3] = {100.0f * (float)(ctr[0] >> 8) / 16777216.0f,
                          100.0f * (float)(ctr[1] >> 8) / 16777216.0f,
                          100.0f * (float)(ctr[2] >> 8) / 16777216.0f};
    for (size_t i = 0; i < cur.size(); i++) {
        const V3 p = cur[i];
        const float n = fbm3(p.x + off[0], p.y + off[1], p.z + off[2]);
        const float r = 1.0f + amp * (2.0f * n - 1.0f);
        out_v9[3 * i] = p.x * r;
        out_v9[3 * i + 1] = p.y * r;
        out_v9[3 * i + 2] = p.z * r;
    }
    return RG_OK;
}
