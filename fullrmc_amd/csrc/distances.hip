// distances.hip -- pairs_distances entry points on gfx950.
//
// One kernel family computes "point vs all atoms" distances or differences
// for k points at once (grid.y = point); the point is either an atom of the
// array (index mode) or an explicit (3,k) column (point mode).  Each thread
// handles one atom row; loads of the (n,3) coordinate rows are contiguous
// across the wave.  The basis is read from device memory by every thread
// (scalar loads, uniform), so no host synchronisation is needed per call.
#include "frmc_device.h"
#include "frmc_host.h"

namespace frmc {

enum PointMode { PM_INDEX = 0, PM_POINTS = 1 };

__device__ __forceinline__ void load_basis_dev(const float *__restrict__ basis, Basis &B) {
#pragma unroll
    for (int k = 0; k < 9; k++) B.b[k] = basis[k];
}

// out layout: distances (n,k) -> out[i*k+c]; differences (n,3,k) -> out[(i*3+d)*k+c]
template <bool PBC, bool DIFF>
__global__ void __launch_bounds__(256) k_vs_all(const float *__restrict__ coords, int32_t n,
                                                const float *__restrict__ basis, int mode,
                                                const int32_t *__restrict__ indexes, int32_t atomIndex,
                                                const float *__restrict__ points, int32_t k, int allAtoms,
                                                float *__restrict__ out) {
    const int32_t c = blockIdx.y;
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float px, py, pz;
    int32_t start = 0;
    if (mode == PM_INDEX) {
        int32_t a = indexes ? indexes[c] : atomIndex;
        if (a < 0 || a >= n) return;  // out-of-range index: nothing written (host shim rejects it)
        px = coords[3 * (int64_t)a];
        py = coords[3 * (int64_t)a + 1];
        pz = coords[3 * (int64_t)a + 2];
        start = allAtoms ? 0 : a;
    } else {
        px = points[c];
        py = points[k + c];
        pz = points[2 * k + c];
    }
    if (i < start) return;  // reference leaves [0,atomIndex) uninitialised
    const float cx = coords[3 * (int64_t)i], cy = coords[3 * (int64_t)i + 1], cz = coords[3 * (int64_t)i + 2];
    // pairs_distances.pyx:373-375 / 461-463: diff = point - coords[i]
    float dx = px - cx, dy = py - cy, dz = pz - cz;
    float rx, ry, rz;
    if (PBC) {
        Basis B;
        load_basis_dev(basis, B);
        pbc_real(dx, dy, dz, B, rx, ry, rz);
    } else {
        rx = dx;
        ry = dy;
        rz = dz;
    }
    if (DIFF) {
        out[((int64_t)i * 3 + 0) * k + c] = rx;
        out[((int64_t)i * 3 + 1) * k + c] = ry;
        out[((int64_t)i * 3 + 2) * k + c] = rz;
    } else {
        out[(int64_t)i * k + c] = cr_sqrt(norm2(rx, ry, rz));
    }
}

// Row-wise pairs: from_to_points_differences (diff = to[i]-from[i]),
// pair_difference_to_point (n=1, diff = p2-p1), point_to_point_distance
// (n=1, diff = p1-p2, distance).
template <bool PBC, bool DIFF>
__global__ void __launch_bounds__(256) k_rowwise(const float *__restrict__ a, const float *__restrict__ b,
                                                 int32_t n, const float *__restrict__ basis,
                                                 float *__restrict__ out) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float dx = a[3 * (int64_t)i] - b[3 * (int64_t)i];
    float dy = a[3 * (int64_t)i + 1] - b[3 * (int64_t)i + 1];
    float dz = a[3 * (int64_t)i + 2] - b[3 * (int64_t)i + 2];
    float rx, ry, rz;
    if (PBC) {
        Basis B;
        load_basis_dev(basis, B);
        pbc_real(dx, dy, dz, B, rx, ry, rz);
    } else {
        rx = dx;
        ry = dy;
        rz = dz;
    }
    if (DIFF) {
        out[3 * (int64_t)i] = rx;
        out[3 * (int64_t)i + 1] = ry;
        out[3 * (int64_t)i + 2] = rz;
    } else {
        out[i] = cr_sqrt(norm2(rx, ry, rz));
    }
}

static int launch_vs_all(bool diff, const float *coords, int32_t n, const float *basis, int isPBC, int mode,
                         const int32_t *indexes, int32_t atomIndex, const float *points, int32_t k, int allAtoms,
                         float *out, void *stream) {
    FRMC_REQUIRE(n >= 0 && k >= 0, "negative size (n=%d, k=%d)", n, k);
    FRMC_REQUIRE(coords != nullptr || n == 0, "coords is NULL");
    FRMC_REQUIRE(out != nullptr || n == 0 || k == 0, "out is NULL");
    FRMC_REQUIRE(!isPBC || basis != nullptr, "basis is NULL");
    if (n == 0 || k == 0) return FRMC_OK;
    FRMC_REQUIRE(k <= 65535, "too many points (k=%d > 65535)", k);
    dim3 grid(blocks_for(n, 256), (unsigned)k);
    hipStream_t st = as_stream(stream);
    if (isPBC) {
        if (diff)
            k_vs_all<true, true><<<grid, 256, 0, st>>>(coords, n, basis, mode, indexes, atomIndex, points, k,
                                                       allAtoms, out);
        else
            k_vs_all<true, false><<<grid, 256, 0, st>>>(coords, n, basis, mode, indexes, atomIndex, points, k,
                                                        allAtoms, out);
    } else {
        if (diff)
            k_vs_all<false, true><<<grid, 256, 0, st>>>(coords, n, basis, mode, indexes, atomIndex, points, k,
                                                        allAtoms, out);
        else
            k_vs_all<false, false><<<grid, 256, 0, st>>>(coords, n, basis, mode, indexes, atomIndex, points,
                                                         k, allAtoms, out);
    }
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

static int launch_rowwise(bool diff, const float *a, const float *b, int32_t n, const float *basis, int isPBC,
                          float *out, void *stream) {
    FRMC_REQUIRE(n >= 0, "negative size");
    FRMC_REQUIRE(!isPBC || basis != nullptr, "basis is NULL");
    if (n == 0) return FRMC_OK;
    FRMC_REQUIRE(a && b && out, "NULL pointer argument");
    hipStream_t st = as_stream(stream);
    unsigned g = blocks_for(n, 256);
    if (isPBC) {
        if (diff) k_rowwise<true, true><<<g, 256, 0, st>>>(a, b, n, basis, out);
        else      k_rowwise<true, false><<<g, 256, 0, st>>>(a, b, n, basis, out);
    } else {
        if (diff) k_rowwise<false, true><<<g, 256, 0, st>>>(a, b, n, basis, out);
        else      k_rowwise<false, false><<<g, 256, 0, st>>>(a, b, n, basis, out);
    }
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

}  // namespace frmc

using namespace frmc;

extern "C" {

int frmc_pairs_distances_to_indexcoords(int32_t atomIndex, const float *coords, int32_t n, const float *basis,
                                        int isPBC, int allAtoms, float *out, void *stream) {
    FRMC_REQUIRE(atomIndex >= 0 && atomIndex < n, "atomIndex %d out of range [0,%d)", atomIndex, n);
    return launch_vs_all(false, coords, n, basis, isPBC, PM_INDEX, nullptr, atomIndex, nullptr, 1, allAtoms, out,
                         stream);
}

int frmc_pairs_distances_to_multi_indexcoords(const int32_t *indexes, int32_t k, const float *coords, int32_t n,
                                              const float *basis, int isPBC, int allAtoms, float *out,
                                              void *stream) {
    FRMC_REQUIRE(indexes != nullptr || k == 0, "indexes is NULL");
    return launch_vs_all(false, coords, n, basis, isPBC, PM_INDEX, indexes, 0, nullptr, k, allAtoms, out, stream);
}

int frmc_pairs_distances_to_point(const float *point, const float *coords, int32_t n, const float *basis,
                                  int isPBC, float *out, void *stream) {
    FRMC_REQUIRE(point != nullptr, "point is NULL");
    return launch_vs_all(false, coords, n, basis, isPBC, PM_POINTS, nullptr, 0, point, 1, 1, out, stream);
}

int frmc_pairs_distances_to_multi_points(const float *points, int32_t k, const float *coords, int32_t n,
                                         const float *basis, int isPBC, float *out, void *stream) {
    FRMC_REQUIRE(points != nullptr || k == 0, "points is NULL");
    return launch_vs_all(false, coords, n, basis, isPBC, PM_POINTS, nullptr, 0, points, k, 1, out, stream);
}

int frmc_point_to_point_distance(const float *point1, const float *point2, const float *basis, int isPBC,
                                 float *out, void *stream) {
    // pairs_distances.pyx:256-282: diff = point1 - point2
    return launch_rowwise(false, point1, point2, 1, basis, isPBC, out, stream);
}

int frmc_pairs_differences_to_indexcoords(int32_t atomIndex, const float *coords, int32_t n, const float *basis,
                                          int isPBC, int allAtoms, float *out, void *stream) {
    FRMC_REQUIRE(atomIndex >= 0 && atomIndex < n, "atomIndex %d out of range [0,%d)", atomIndex, n);
    return launch_vs_all(true, coords, n, basis, isPBC, PM_INDEX, nullptr, atomIndex, nullptr, 1, allAtoms, out,
                         stream);
}

int frmc_pairs_differences_to_multi_indexcoords(const int32_t *indexes, int32_t k, const float *coords,
                                                int32_t n, const float *basis, int isPBC, int allAtoms,
                                                float *out, void *stream) {
    FRMC_REQUIRE(indexes != nullptr || k == 0, "indexes is NULL");
    return launch_vs_all(true, coords, n, basis, isPBC, PM_INDEX, indexes, 0, nullptr, k, allAtoms, out, stream);
}

int frmc_pairs_differences_to_point(const float *point, const float *coords, int32_t n, const float *basis,
                                    int isPBC, float *out, void *stream) {
    FRMC_REQUIRE(point != nullptr, "point is NULL");
    return launch_vs_all(true, coords, n, basis, isPBC, PM_POINTS, nullptr, 0, point, 1, 1, out, stream);
}

int frmc_pairs_differences_to_multi_points(const float *points, int32_t k, const float *coords, int32_t n,
                                           const float *basis, int isPBC, float *out, void *stream) {
    FRMC_REQUIRE(points != nullptr || k == 0, "points is NULL");
    return launch_vs_all(true, coords, n, basis, isPBC, PM_POINTS, nullptr, 0, points, k, 1, out, stream);
}

int frmc_from_to_points_differences(const float *pointsFrom, const float *pointsTo, int32_t n,
                                    const float *basis, int isPBC, float *out, void *stream) {
    // pairs_distances.pyx:84-93 / 127-132: diff = to[i] - from[i]
    return launch_rowwise(true, pointsTo, pointsFrom, n, basis, isPBC, out, stream);
}

int frmc_pair_difference_to_point(const float *point1, const float *point2, const float *basis, int isPBC,
                                  float *out, void *stream) {
    // pairs_distances.pyx:54-59 / 108-114: diff = point2 - point1
    return launch_rowwise(true, point2, point1, 1, basis, isPBC, out, stream);
}

}  // extern "C"
