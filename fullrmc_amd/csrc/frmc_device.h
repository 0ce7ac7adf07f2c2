// frmc_device.h -- device-side float32 arithmetic of the pair path.
//
// Every helper reproduces the IEEE binary32 expression tree of the
// reference Cython loops (left-to-right, round-to-nearest-even, no fused
// multiply-add: the library is compiled with -ffp-contract=off), so a GPU
// distance or bin is bit-identical to the one the reference computes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace frmc {

// Uniform (wave-invariant) loads of data that an earlier kernel or copy
// wrote go through coherent vector loads: a plain load from a uniform address
// is compiled to a scalar load through the scalar data cache, which has
// returned stale values here between back-to-back kernels of a stream.
// Lane-varying loads use the vector path and need no such care.
template <typename T>
__device__ __forceinline__ T load_dev(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// pairs_distances.pyx:31-32: x>0 ? floor(x+0.5) : ceil(x-0.5).  floor/ceil of
// a float are exact, so the float versions equal libm's double path.
__device__ __forceinline__ float ref_round(float x) {
    return x > 0.0f ? floorf(x + 0.5f) : ceilf(x - 0.5f);
}

// b = x - ref_round(x) in four operations: x - copysign(floor(|x| + 0.5f), x).
// For x > 0 it is the reference's expression.  For x < 0, x - 0.5f rounds to
// exactly -(|x| + 0.5f) (round-to-nearest-even is symmetric under negation)
// and ceil(-y) = -floor(y), so the rounded value is the same, signed zero
// included; for x = +-0 both forms give b = +0.
__device__ __forceinline__ float min_image(float x) { return x - copysignf(floorf(fabsf(x) + 0.5f), x); }

// |x| - floor(|x| + 0.5f): equal to +-min_image(x) (for x < 0,
// x + f = -(|x| - f) exactly), for callers that only square it.
__device__ __forceinline__ float min_image_pm(float x) {
    const float a = fabsf(x);
    return a - floorf(a + 0.5f);
}

struct Basis {
    float b[9];  // row-major 3x3, b[3*r+c] = basis[r,c]
};

// Minimum-image real-space difference for a fractional difference
// (pairs_distances.pyx:373-382): b = d - round(d); r_c = (bx*B0c + by*B1c) + bz*B2c.
__device__ __forceinline__ void pbc_real(float dx, float dy, float dz, const Basis &B, float &rx, float &ry,
                                         float &rz) {
    float bx = min_image(dx);
    float by = min_image(dy);
    float bz = min_image(dz);
    rx = bx * B.b[0] + by * B.b[3] + bz * B.b[6];
    ry = bx * B.b[1] + by * B.b[4] + bz * B.b[7];
    rz = bx * B.b[2] + by * B.b[5] + bz * B.b[8];
}

// Orthorhombic specialisation for squared norms only: with zero off-diagonal
// basis entries the two extra terms of each component are +-0 and leave a
// non-zero product unchanged (a zero product only changes sign), and the
// components may carry the opposite sign (min_image_pm), so only the squared
// norm of the result is bit-identical to the general form.
__device__ __forceinline__ void pbc_real_ortho(float dx, float dy, float dz, const Basis &B, float &rx,
                                               float &ry, float &rz) {
    float bx = min_image_pm(dx);
    float by = min_image_pm(dy);
    float bz = min_image_pm(dz);
    rx = bx * B.b[0];
    ry = by * B.b[4];
    rz = bz * B.b[8];
}

// ((rx*rx + ry*ry) + rz*rz), the float argument of the reference's sqrt
__device__ __forceinline__ float norm2(float rx, float ry, float rz) {
    return rx * rx + ry * ry + rz * rz;
}

// Squared norm of the difference a - c (box coordinates for PBC, real
// coordinates otherwise) as the reference computes it before its sqrt.
template <bool PBC, bool ORTHO>
__device__ __forceinline__ float pair_s(float ax, float ay, float az, float cx, float cy, float cz,
                                        const Basis &B) {
    float dx = ax - cx, dy = ay - cy, dz = az - cz;
    float rx, ry, rz;
    if (!PBC) {
        rx = dx; ry = dy; rz = dz;
    } else if (ORTHO) {
        pbc_real_ortho(dx, dy, dz, B, rx, ry, rz);
    } else {
        pbc_real(dx, dy, dz, B, rx, ry, rz);
    }
    return norm2(rx, ry, rz);
}

// (float)sqrt((double)s) == correctly rounded float sqrt.  The library is
// built with correctly rounded fp32 divide/sqrt (hipcc default), which
// __builtin_sqrtf honours; a device known-answer test checks it.
__device__ __forceinline__ float cr_sqrt(float s) { return __builtin_sqrtf(s); }
__device__ __forceinline__ float cr_div(float a, float b) { return a / b; }

// Histogram grid for one constraint.  s_lo/s_hi are the exact squared-norm
// thresholds equivalent to the reference's distance tests
// (d<minD rejected  <=>  s < s_lo;  d>=maxD rejected  <=>  s >= s_hi), both
// computed on the host from the monotonicity of correctly rounded sqrt.
struct Grid {
    float minD, maxD, bin;
    float s_lo, s_hi;
    int32_t H;
};

// Bin index of a squared norm already known to be inside [s_lo, s_hi):
// pairs_histograms.pyx:64  binIndex = (int)((distance-minDistance)/bin)
__device__ __forceinline__ int32_t bin_of(float s, const Grid &g) {
    float d = cr_sqrt(s);
    return (int32_t)cr_div(d - g.minD, g.bin);
}

__device__ __forceinline__ bool in_range(float s, const Grid &g) { return s >= g.s_lo && s < g.s_hi; }

struct HistArgs {
    Grid g;
    int32_t nel;
    int64_t size;  // nel*nel*H
};

}  // namespace frmc
