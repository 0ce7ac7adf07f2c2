// common.hip -- error state, exact squared-norm thresholds, and the
// boundary_conditions_collection entry points.
#include <atomic>
#include <math.h>
#include <stdarg.h>

#include "frmc_device.h"
#include "frmc_host.h"

namespace frmc {

static thread_local char g_err[1024] = "";
static thread_local int64_t g_overflow = 0;
static thread_local int g_timing = 0;
static thread_local float g_last_ms = 0.f;

bool timing_enabled() { return g_timing != 0; }
void set_last_kernel_ms(float ms) { g_last_ms = ms; }

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

void set_overflow(int64_t v) { g_overflow = v; }

static inline float host_cr_sqrt(float s) {
    // binary64 sqrt rounded to binary32 is the correctly rounded binary32
    // sqrt (53 >= 2*24+2), which is what the reference's (float)sqrt(double) is.
    return (float)sqrt((double)s);
}

float sq_threshold(float t) {
    if (!(t > 0.0f)) return -INFINITY;  // every s >= 0 satisfies sqrt(s) >= t
    if (isinf(t)) return INFINITY;
    // binary search over the ordered bit patterns of non-negative floats for
    // the first s with host_cr_sqrt(s) >= t  (monotone in s)
    uint32_t lo = 0, hi = 0x7f800000u;  // +0 .. +inf ; sqrt(+inf) = inf >= t
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        float s;
        memcpy(&s, &mid, 4);
        if (host_cr_sqrt(s) >= t) hi = mid; else lo = mid + 1;
    }
    float r;
    memcpy(&r, &lo, 4);
    return r;
}

hipError_t pool_malloc(void **p, size_t bytes, hipStream_t st) {
    static std::atomic<uint64_t> retained{0};  // bit d: device d's pool configured
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 64 && !(retained.load() & (1ull << dev))) {
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
            uint64_t keep = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
        retained.fetch_or(1ull << dev);
    }
    return hipMallocAsync(p, bytes, st);
}

int load_basis(const float *dbasis, Basis &B, bool &ortho, hipStream_t st) {
    FRMC_CHECK_HIP(hipMemcpyAsync(B.b, dbasis, sizeof(B.b), hipMemcpyDeviceToHost, st));
    FRMC_CHECK_HIP(hipStreamSynchronize(st));
    const float *b = B.b;
    ortho = b[1] == 0.f && b[2] == 0.f && b[3] == 0.f && b[5] == 0.f && b[6] == 0.f && b[7] == 0.f;
    return FRMC_OK;
}

Grid make_grid(float minD, float maxD, float bin, int32_t H) {
    Grid g;
    g.minD = minD;
    g.maxD = maxD;
    g.bin = bin;
    g.H = H;
    g.s_lo = sq_threshold(minD);
    g.s_hi = sq_threshold(maxD);
    return g;
}

// get_reciprocal_basis (boundary_conditions_collection.pyx:45-80), one thread.
__global__ void k_reciprocal_basis(const float *__restrict__ b, float *__restrict__ r, double *__restrict__ vol) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float c[9];
    c[0] = b[4] * b[8] - b[7] * b[5];
    c[1] = -(b[1] * b[8] - b[7] * b[1]);  // the reference reuses basis[0,1] here (:59)
    c[2] = b[1] * b[5] - b[4] * b[2];
    c[3] = -(b[3] * b[8] - b[6] * b[5]);
    c[4] = b[0] * b[8] - b[6] * b[2];
    c[5] = -(b[0] * b[5] - b[3] * b[2]);
    c[6] = b[3] * b[7] - b[6] * b[4];
    c[7] = -(b[0] * b[7] - b[6] * b[1]);
    c[8] = b[0] * b[4] - b[3] * b[1];
    // vol is a C double holding the float32 triple product; the divisions run
    // in double and round to float32 when stored.
    double v = (double)(b[0] * c[0] + b[3] * c[1] + b[6] * c[2]);
    for (int k = 0; k < 9; k++) r[k] = (float)((double)c[k] / v);
    *vol = v;
}

// transform_coordinates (boundary_conditions_collection.pyx:105-108)
__global__ void __launch_bounds__(256) k_transform(const float *__restrict__ M, const float *__restrict__ c,
                                                   int32_t n, float *__restrict__ out) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = c[3 * (int64_t)i], y = c[3 * (int64_t)i + 1], z = c[3 * (int64_t)i + 2];
    out[3 * (int64_t)i] = x * M[0] + y * M[3] + z * M[6];
    out[3 * (int64_t)i + 1] = x * M[1] + y * M[4] + z * M[7];
    out[3 * (int64_t)i + 2] = x * M[2] + y * M[5] + z * M[8];
}

// box_coordinates_real_distances (boundary_conditions_collection.pyx:150-163):
// diff = boxCoords[i] - boxCoords[atomIndex]
__global__ void __launch_bounds__(256) k_box_real_distances(int32_t atomIndex, const int32_t *__restrict__ idx,
                                                            int32_t k, int32_t n, const float *__restrict__ c,
                                                            const float *__restrict__ basis,
                                                            float *__restrict__ out) {
    const int32_t ii = blockIdx.x * blockDim.x + threadIdx.x;
    if (ii >= k) return;
    Basis B;
#pragma unroll
    for (int q = 0; q < 9; q++) B.b[q] = basis[q];
    const int64_t i = idx[ii];
    if (i < 0 || i >= n) {  // guarded; the host shim rejects such lists
        out[ii] = __int_as_float(0x7fc00000);
        return;
    }
    float dx = c[3 * i] - c[3 * (int64_t)atomIndex];
    float dy = c[3 * i + 1] - c[3 * (int64_t)atomIndex + 1];
    float dz = c[3 * i + 2] - c[3 * (int64_t)atomIndex + 2];
    float rx, ry, rz;
    pbc_real(dx, dy, dz, B, rx, ry, rz);
    out[ii] = cr_sqrt(norm2(rx, ry, rz));
}

}  // namespace frmc

using namespace frmc;

extern "C" {

const char *frmc_last_error(void) { return g_err; }
int frmc_version(void) { return 10000; }
int64_t frmc_last_overflow(void) { return g_overflow; }
void frmc_set_timing(int enable) { g_timing = enable; }
float frmc_last_kernel_ms(void) { return g_last_ms; }

int frmc_get_reciprocal_basis(const float *basis, float *rbasis, double *vol, void *stream) {
    FRMC_REQUIRE(basis && rbasis && vol, "NULL pointer argument");
    hipStream_t st = as_stream(stream);
    double *dvol = nullptr;
    FRMC_CHECK_HIP(pool_malloc((void **)&dvol, sizeof(double), st));
    k_reciprocal_basis<<<1, 64, 0, st>>>(basis, rbasis, dvol);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(vol, dvol, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFreeAsync(dvol, st);
    if (e != hipSuccess) {
        set_error("get_reciprocal_basis failed: %s", hipGetErrorString(e));
        return FRMC_EHIP;
    }
    return FRMC_OK;
}

int frmc_transform_coordinates(const float *M, const float *coords, int32_t n, float *out, void *stream) {
    FRMC_REQUIRE(n >= 0, "negative size");
    if (n == 0) return FRMC_OK;
    FRMC_REQUIRE(M && coords && out, "NULL pointer argument");
    k_transform<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(M, coords, n, out);
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

int frmc_box_coordinates_real_distances(int32_t atomIndex, const int32_t *indexes, int32_t k, const float *coords,
                                        int32_t n, const float *basis, float *out, void *stream) {
    FRMC_REQUIRE(k >= 0 && n >= 0, "negative size");
    FRMC_REQUIRE(atomIndex >= 0 && atomIndex < n, "atomIndex %d out of range [0,%d)", atomIndex, n);
    if (k == 0) return FRMC_OK;
    FRMC_REQUIRE(indexes && coords && basis && out, "NULL pointer argument");
    k_box_real_distances<<<blocks_for(k, 256), 256, 0, as_stream(stream)>>>(atomIndex, indexes, k, n, coords,
                                                                             basis, out);
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

}  // extern "C"
