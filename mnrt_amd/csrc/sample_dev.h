// sample_dev.h — counter-based random numbers and trig-free direction sampling shared
// by the photon tracer and the gather-ray generator.
//
// Random numbers: the lowbias32 integer finaliser applied to (seed, path, bounce, draw).
// Cosine-distributed directions: Malley's method with a rejection-sampled point of the
// unit disk and the orthonormal basis of Duff et al. 2017 ("Building an Orthonormal
// Basis, Revisited"), so the sampling needs only +, *, / and sqrt.  (The reference draws
// Mersenne-Twister uniforms and maps them with sin/cos, sample_dev.h:100-279.)
#pragma once

#include "trace_dev.h"

namespace mnrt {

__host__ __device__ __forceinline__ uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ float rnd01(uint32_t seed, uint32_t path, uint32_t bounce, uint32_t draw) {
    uint32_t h = hash32(seed ^ hash32(path ^ hash32(bounce * 0x9E3779B9u ^ hash32(draw))));
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}

// Cosine-distributed direction about unit normal n.
__device__ __forceinline__ f3 cosine_dir(f3 n, uint32_t seed, uint32_t path, uint32_t bounce) {
    float x = 0.0f, y = 0.0f, r2 = 0.0f;
    bool ok = false;
    for (uint32_t t = 0; t < 32 && !ok; ++t) {
        x = 2.0f * rnd01(seed, path, bounce, 2 * t + 1) - 1.0f;
        y = 2.0f * rnd01(seed, path, bounce, 2 * t + 2) - 1.0f;
        r2 = x * x + y * y;
        ok = r2 < 1.0f;
    }
    if (!ok) {
        x = 0.0f;
        y = 0.0f;
        r2 = 0.0f;
    }
    float z = sqrtf(1.0f - r2);
    float sign = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    f3 T = mk3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    f3 B = mk3(b, sign + n.y * n.y * a, -n.y);
    return mk3(x * T.x + y * B.x + z * n.x, x * T.y + y * B.y + z * n.y, x * T.z + y * B.z + z * n.z);
}

__device__ __forceinline__ f3 unit_normal(f3 e1, f3 e2) {
    f3 n = cross3(e1, e2);
    float len = sqrtf(dot3(n, n));
    return mk3(n.x / len, n.y / len, n.z / len);
}

}  // namespace mnrt
