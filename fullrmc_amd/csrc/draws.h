// draws.h -- the arithmetic of a translation proposal from MT19937 words,
// shared by the host loop (driver.hip) and the device window loop
// (window_loop.hip), so that both round every step identically.
//
// Both of fullrmc's streams are MT19937 (Matsumoto & Nishimura 1998):
// CPython's random (group index, amplitude, acceptance draw) and NumPy's
// legacy RandomState (direction).  A double in [0, 1) takes two 32-bit words
// (random.random() and RandomState.random_sample() alike).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace frmc {

constexpr int kMtN = 624;  // words of an MT19937 state block

__host__ __device__ inline uint32_t mt_temper(uint32_t y) {
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

// the twist term of word i: upper bit of word i, lower 31 bits of word i+1
__host__ __device__ inline uint32_t mt_mix(uint32_t wi, uint32_t wnext) {
    const uint32_t y = (wi & 0x80000000u) | (wnext & 0x7fffffffu);
    return (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
}

// random.random() / random_sample() from two tempered words
__host__ __device__ inline double mt_uniform(uint32_t w0, uint32_t w1) {
    return ((w0 >> 5) * 67108864.0 + (w1 >> 6)) * (1.0 / 9007199254740992.0);
}

// bits of randrange(n) = getrandbits(n.bit_length()) for 1 <= n < 2^31
__host__ __device__ inline int32_t bit_length(int32_t n) {
    int32_t k = 0;
    while (k < 32 && ((uint32_t)1 << k) <= (uint32_t)n) k++;
    return k;
}

// np.linalg.norm of a float32 3-vector: x.dot(x) goes through BLAS sdot,
// which rounds each product to float32 and accumulates in double; NumPy
// casts the sum to float32 and takes the float32 sqrt.
__host__ __device__ inline float np_norm3(const float v[3]) {
    const float p0 = v[0] * v[0], p1 = v[1] * v[1], p2 = v[2] * v[2];
    double acc = 0.0;
    acc += (double)p0;
    acc += (double)p1;
    acc += (double)p2;
    return sqrtf((float)acc);
}

// One try of generate_random_vector's direction (Core/Collection.py:684-690)
// from three uniforms of the NumPy stream: false when its norm is 0 (the
// reference then draws three more).
__host__ __device__ inline bool random_direction(const double u[3], float v[3], float &norm) {
    for (int d = 0; d < 3; d++) v[d] = (float)(1.0 - 2.0 * u[d]);
    norm = np_norm3(v);
    return norm != 0.0f;
}

// The rest of generate_random_vector (Core/Collection.py:691-701): unit
// direction, amplitude uniform ua of the Python stream in [minAmp, maxAmp).
__host__ __device__ inline void scale_direction(float v[3], float norm, double ua, float minAmp, float maxAmp) {
    for (int d = 0; d < 3; d++) v[d] = v[d] / norm;
    float base[3];
    for (int d = 0; d < 3; d++) base[d] = v[d] * minAmp;
    const float amp = maxAmp - minAmp;
    const float scale = (float)ua * amp;  // Python float * float32 -> float32
    for (int d = 0; d < 3; d++) v[d] = v[d] * scale;
    for (int d = 0; d < 3; d++) v[d] = v[d] + base[d];
}

// transform_coordinates (boundary_conditions_collection.pyx:105-108)
__host__ __device__ inline void to_box(const float M[9], const float r[3], float b[3]) {
    b[0] = r[0] * M[0] + r[1] * M[3] + r[2] * M[6];
    b[1] = r[0] * M[1] + r[1] * M[4] + r[2] * M[7];
    b[2] = r[0] * M[2] + r[1] * M[5] + r[2] * M[8];
}

}  // namespace frmc
