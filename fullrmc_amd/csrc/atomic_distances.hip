// atomic_distances.hip -- the atomic_distances extension
// (Extensions/atomic_distances.pyx) on gfx950: per element pair, the number
// and the float32 sum of the (reduced) distances of atom pairs that pass the
// molecule and distance-limit filters, as InterMolecularDistanceConstraint
// and IntraMolecularDistanceConstraint use them.
//
// The reference adds each selected distance to its slot in a fixed order
// (atoms in `indexes` order, then other atoms by index), and float32
// addition is not associative, so the sums are rebuilt in that order:
//   k_ad_pass<EMIT=false>  every (atom, other atom) pair in parallel:
//                distance, filters; per-slot counts (LDS, then global atomics)
//   k_ad_pass<EMIT=true>   the same pass writes each selected pair as a
//                64-bit key (slot, atom position, other index) and its value
//   radix sort   rocPRIM, by key: each slot's terms become contiguous and in
//                the reference's order
//   k_ad_sum     one lane per slot adds its terms sequentially onto the
//                slot's existing float32 value and adds the count
// Counts are integers and need no order.  The first pass's counts size the
// second pass (one host synchronisation per call).
#include <rocprim/device/device_radix_sort.hpp>
#include <vector>

#include "frmc_device.h"
#include "frmc_host.h"

namespace frmc {

namespace {

enum AdSource { AD_PBC = 0, AD_IBC = 1, AD_DISTS = 2 };

struct AdArgs {
    const float *coords;     // (n,3) box (PBC) or real (IBC) coordinates, or NULL
    const float *dists;      // (n,k) distances, or NULL
    int32_t n, k;            // atoms (rows), atoms in the index list (columns of dists)
    Basis B;
    const int32_t *indexes;  // [k] atom of each column, or NULL: atomIndex
    int32_t atomIndex;
    const int32_t *mol, *elem;
    int32_t nel;
    const float *lower, *upper;  // [nel][nel], read as [other][atom]
    int32_t flags, allAtoms;
};

// distance of atom a (row of the index list) to atom i, as pairs_distances
// computes it (pairs_distances.pyx:343-386 / :438-471: diff = coords[a] - coords[i])
template <int SRC>
__device__ __forceinline__ float ad_distance(const AdArgs &A, int32_t c, int32_t a, float px, float py, float pz,
                                             int32_t i) {
    if (SRC == AD_DISTS) return A.dists[(int64_t)i * A.k + c];
    const float dx = px - A.coords[3 * (int64_t)i], dy = py - A.coords[3 * (int64_t)i + 1],
                dz = pz - A.coords[3 * (int64_t)i + 2];
    float rx = dx, ry = dy, rz = dz;
    if (SRC == AD_PBC) pbc_real(dx, dy, dz, A.B, rx, ry, rz);
    return cr_sqrt(norm2(rx, ry, rz));
}

// _single_atomic_distances_dists (atomic_distances.pyx:42-118) for one pair:
// true if counted, with the counted value and the slot
// ws = (0 intra | 1 inter)*nel*nel + e_atom*nel + e_other.
__device__ __forceinline__ bool ad_select(const AdArgs &A, float d, int32_t i, int32_t a, int32_t ma, int32_t ea,
                                          float &v, int32_t &ws) {
    if (i == a) return false;
    const int32_t mi = A.mol[i];
    const bool same = mi == ma;
    if (!(A.flags & FRMC_AD_INTRA) && same) return false;
    if (!(A.flags & FRMC_AD_INTER) && !same) return false;
    const int32_t ei = A.elem[i];
    const float lower = A.lower[ei * A.nel + ea], upper = A.upper[ei * A.nel + ea];
    if (A.flags & FRMC_AD_WITHIN) {
        if (d < lower) return false;
        if (d >= upper) return false;
    } else if (d >= lower && d < upper) {
        return false;
    }
    if (A.flags & FRMC_AD_RED_UPPER) {
        d = fabsf(upper - d);
    } else if (A.flags & FRMC_AD_RED_LOWER) {
        d = fabsf(lower - d);
    } else if (A.flags & FRMC_AD_RED) {
        const float mid = (lower + upper) / 2.0f;
        d = d > mid ? fabsf(upper - d) : fabsf(lower - d);
    }
    v = d;
    ws = (same ? 0 : A.nel * A.nel) + ea * A.nel + ei;
    return true;
}

// grid (ceil(n/256), k): column c = atom indexes[c], one other atom per thread.
// EMIT=false: per-slot counts into cnt (LDS first).  EMIT=true: each
// selected pair to (keys, vals) at an atomic position.
template <int SRC, bool EMIT>
__global__ void __launch_bounds__(256) k_ad_pass(AdArgs A, uint32_t *__restrict__ cnt, uint32_t *__restrict__ fill,
                                                 uint64_t *__restrict__ keys, float *__restrict__ vals) {
    extern __shared__ uint32_t lcnt[];  // [2*nel*nel] (count pass)
    const int32_t c = blockIdx.y;
    const int32_t nws = 2 * A.nel * A.nel;
    if (!EMIT) {
        for (int32_t s = threadIdx.x; s < nws; s += blockDim.x) lcnt[s] = 0;
        __syncthreads();
    }
    int32_t a = A.indexes ? load_dev(A.indexes + c) : A.atomIndex;
    const bool aok = a >= 0 && a < A.n;  // the host shims reject other indexes; nothing is counted for them
    a = aok ? a : 0;
    const int32_t ma = load_dev(A.mol + a), ea = load_dev(A.elem + a);
    float px = 0.f, py = 0.f, pz = 0.f;
    if (SRC != AD_DISTS) {
        px = load_dev(A.coords + 3 * (int64_t)a);
        py = load_dev(A.coords + 3 * (int64_t)a + 1);
        pz = load_dev(A.coords + 3 * (int64_t)a + 2);
    }
    const int32_t start = A.allAtoms ? 0 : a;
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float v = 0.f;
    int32_t ws = 0;
    bool sel = false;
    if (aok && i >= start && i < A.n) {
        const float d = ad_distance<SRC>(A, c, a, px, py, pz, i);
        sel = ad_select(A, d, i, a, ma, ea, v, ws);
    }
    if (EMIT) {
        if (sel) {
            const uint32_t pos = atomicAdd(fill, 1u);
            keys[pos] = ((uint64_t)ws << 40) | ((uint64_t)c * (uint64_t)A.n + (uint64_t)i);
            vals[pos] = v;
        }
    } else {
        if (sel) atomicAdd(&lcnt[ws], 1u);
        __syncthreads();
        for (int32_t s = threadIdx.x; s < nws; s += blockDim.x)
            if (lcnt[s]) atomicAdd(&cnt[s], lcnt[s]);
    }
}

// one lane per slot: dsum[slot] += terms in order (float32), nsum[slot] += count
__global__ void __launch_bounds__(256) k_ad_sum(const uint32_t *__restrict__ cnt, int32_t nws,
                                                const float *__restrict__ vals, float *__restrict__ dintra,
                                                float *__restrict__ dinter, int32_t *__restrict__ nintra,
                                                int32_t *__restrict__ ninter) {
    const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nws) return;
    uint64_t off = 0;
    for (int32_t t = 0; t < s; t++) off += cnt[t];
    const uint32_t len = cnt[s];
    const int32_t half = nws / 2;
    float *d = s < half ? dintra + s : dinter + (s - half);
    int32_t *nn = s < half ? nintra + s : ninter + (s - half);
    float acc = *d;
    for (uint32_t j = 0; j < len; j++) acc = acc + vals[off + j];
    *d = acc;
    *nn = *nn + (int32_t)len;
}

// pair_elements_stats (atomic_distances.pyx:632-656): every pair i<j once,
// nintra/ninter[e_i][e_j] += 1 by molecule.  Grid (row tiles of 256,
// column chunks of kPeChunk); LDS-private counters.
constexpr int kPeChunk = 2048;

__global__ void __launch_bounds__(256) k_pair_elements_stats(const int32_t *__restrict__ elem,
                                                             const int32_t *__restrict__ mol, int32_t n, int32_t nel,
                                                             int32_t *__restrict__ nintra,
                                                             int32_t *__restrict__ ninter) {
    extern __shared__ int32_t sm[];
    const int32_t nn = nel * nel;
    int32_t *lc = sm;                  // [2*nn]
    int32_t *ce = sm + 2 * nn;         // [kPeChunk] chunk elements
    int32_t *cm = ce + kPeChunk;       // [kPeChunk] chunk molecules
    const int32_t r0 = blockIdx.x * 256, j0 = blockIdx.y * kPeChunk;
    if (j0 + kPeChunk <= r0 + 1) return;  // chunk entirely at or below the row tile's diagonal (uniform)
    for (int32_t s = threadIdx.x; s < 2 * nn; s += blockDim.x) lc[s] = 0;
    const int32_t len = min(kPeChunk, n - j0);
    for (int32_t t = threadIdx.x; t < len; t += blockDim.x) {
        ce[t] = elem[j0 + t];
        cm[t] = mol[j0 + t];
    }
    __syncthreads();
    const int32_t i = r0 + threadIdx.x;
    if (i < n) {
        const int32_t ei = elem[i], mi = mol[i];
        for (int32_t t = max(0, i + 1 - j0); t < len; t++) {
            const int32_t slot = (cm[t] == mi ? 0 : nn) + ei * nel + ce[t];
            atomicAdd(&lc[slot], 1);
        }
    }
    __syncthreads();
    for (int32_t s = threadIdx.x; s < 2 * nn; s += blockDim.x)
        if (lc[s]) atomicAdd(s < nn ? &nintra[s] : &ninter[s - nn], lc[s]);
}

// Scratch of one call, returned to the stream-ordered pool at scope exit.
struct AdScratch {
    hipStream_t st = nullptr;
    void *p[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    ~AdScratch() {
        for (void *q : p)
            if (q) (void)hipFreeAsync(q, st);
    }
};

template <int SRC>
int ad_run(AdArgs A, float *dintra, float *dinter, int32_t *nintra, int32_t *ninter, hipStream_t st) {
    const int32_t nws = 2 * A.nel * A.nel;
    const size_t lds = sizeof(uint32_t) * nws;
    FRMC_REQUIRE(lds <= 64 * 1024, "atomic_distances: %d element types is too many", A.nel);
    FRMC_REQUIRE((uint64_t)A.k * (uint64_t)A.n < ((uint64_t)1 << 40), "atomic_distances: too many pairs");
    FRMC_REQUIRE(A.k <= 65535, "atomic_distances: at most 65535 atoms per call");
    if (A.k == 0 || A.n == 0) return FRMC_OK;
    AdScratch S;
    S.st = st;
    uint32_t *cnt = nullptr;
    FRMC_CHECK_HIP(pool_malloc(&S.p[0], sizeof(uint32_t) * (nws + 1), st));
    cnt = (uint32_t *)S.p[0];
    FRMC_CHECK_HIP(hipMemsetAsync(cnt, 0, sizeof(uint32_t) * (nws + 1), st));
    const dim3 grid(blocks_for(A.n, 256), A.k);
    hipLaunchKernelGGL((k_ad_pass<SRC, false>), grid, dim3(256), lds, st, A, cnt, nullptr, nullptr, nullptr);
    FRMC_CHECK_LAUNCH();
    std::vector<uint32_t> h(nws);
    FRMC_CHECK_HIP(hipMemcpyAsync(h.data(), cnt, sizeof(uint32_t) * nws, hipMemcpyDeviceToHost, st));
    FRMC_CHECK_HIP(hipStreamSynchronize(st));
    uint64_t L = 0;
    for (uint32_t c : h) L += c;
    FRMC_REQUIRE(L < ((uint64_t)1 << 28), "atomic_distances: %llu counted pairs exceed the 2^28 limit",
                 (unsigned long long)L);
    float *vals_s = nullptr;
    if (L > 0) {
        FRMC_CHECK_HIP(pool_malloc(&S.p[1], sizeof(uint64_t) * L, st));
        FRMC_CHECK_HIP(pool_malloc(&S.p[2], sizeof(uint64_t) * L, st));
        FRMC_CHECK_HIP(pool_malloc(&S.p[3], sizeof(float) * L, st));
        FRMC_CHECK_HIP(pool_malloc(&S.p[4], sizeof(float) * L, st));
        uint64_t *keys = (uint64_t *)S.p[1], *keys_s = (uint64_t *)S.p[2];
        float *vals = (float *)S.p[3];
        vals_s = (float *)S.p[4];
        hipLaunchKernelGGL((k_ad_pass<SRC, true>), grid, dim3(256), 0, st, A, nullptr, cnt + nws, keys, vals);
        FRMC_CHECK_LAUNCH();
        unsigned end_bit = 40;
        while ((1u << (end_bit - 40)) < (unsigned)nws) end_bit++;
        size_t tmp_bytes = 0;
        FRMC_CHECK_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys, keys_s, vals, vals_s, (size_t)L, 0,
                                                 end_bit, st));
        FRMC_CHECK_HIP(pool_malloc(&S.p[5], tmp_bytes ? tmp_bytes : 1, st));
        FRMC_CHECK_HIP(rocprim::radix_sort_pairs(S.p[5], tmp_bytes, keys, keys_s, vals, vals_s, (size_t)L, 0,
                                                 end_bit, st));
    }
    hipLaunchKernelGGL(k_ad_sum, dim3(blocks_for(nws, 256)), dim3(256), 0, st, cnt, nws, vals_s, dintra, dinter,
                       nintra, ninter);
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

int ad_basis(AdArgs &A, const float *basis, int isPBC, hipStream_t st) {
    if (!isPBC) return FRMC_OK;
    bool ortho = false;
    return load_basis(basis, A.B, ortho, st);
}

}  // namespace

}  // namespace frmc

using namespace frmc;

extern "C" int frmc_single_atomic_distances_dists(int32_t atomIndex, const float *distances, int32_t n,
                                                  const int32_t *moleculeIndex, const int32_t *elementIndex,
                                                  float *dintra, float *dinter, int32_t *nintra, int32_t *ninter,
                                                  int32_t numberOfElements, const float *lowerLimit,
                                                  const float *upperLimit, int flags, int allAtoms, void *stream) {
    FRMC_REQUIRE(numberOfElements >= 1 && n >= 0, "atomic_distances: bad sizes");
    FRMC_REQUIRE(atomIndex >= 0 && atomIndex < n, "atomIndex %d outside [0,%d)", atomIndex, n);
    AdArgs A{};
    A.dists = distances;
    A.n = n;
    A.k = 1;
    A.indexes = nullptr;
    A.atomIndex = atomIndex;
    A.mol = moleculeIndex;
    A.elem = elementIndex;
    A.nel = numberOfElements;
    A.lower = lowerLimit;
    A.upper = upperLimit;
    A.flags = flags;
    A.allAtoms = allAtoms;
    return ad_run<AD_DISTS>(A, dintra, dinter, nintra, ninter, as_stream(stream));
}

extern "C" int frmc_multiple_atomic_distances_coords(const int32_t *indexes, int32_t nidx, const float *boxCoords,
                                                     int32_t n, const float *basis, int isPBC,
                                                     const int32_t *moleculeIndex, const int32_t *elementIndex,
                                                     int32_t numberOfElements, const float *lowerLimit,
                                                     const float *upperLimit, int flags, int allAtoms,
                                                     int32_t *nintra, float *dintra, int32_t *ninter, float *dinter,
                                                     void *stream) {
    FRMC_REQUIRE(numberOfElements >= 1 && n >= 0 && nidx >= 0, "atomic_distances: bad sizes");
    hipStream_t st = as_stream(stream);
    AdArgs A{};
    A.coords = boxCoords;
    A.n = n;
    A.k = nidx;
    A.indexes = indexes;
    A.mol = moleculeIndex;
    A.elem = elementIndex;
    A.nel = numberOfElements;
    A.lower = lowerLimit;
    A.upper = upperLimit;
    A.flags = flags;
    A.allAtoms = allAtoms;
    int rc = ad_basis(A, basis, isPBC, st);
    if (rc) return rc;
    return isPBC ? ad_run<AD_PBC>(A, dintra, dinter, nintra, ninter, st)
                 : ad_run<AD_IBC>(A, dintra, dinter, nintra, ninter, st);
}

extern "C" int frmc_multiple_atomic_distances_dists(const int32_t *indexes, int32_t nidx, const float *distances,
                                                    int32_t n, int32_t k, const int32_t *moleculeIndex,
                                                    const int32_t *elementIndex, int32_t numberOfElements,
                                                    const float *lowerLimit, const float *upperLimit, int flags,
                                                    int allAtoms, int32_t *nintra, float *dintra, int32_t *ninter,
                                                    float *dinter, void *stream) {
    FRMC_REQUIRE(numberOfElements >= 1 && n >= 0 && nidx >= 0 && nidx <= k, "atomic_distances: bad sizes");
    AdArgs A{};
    A.dists = distances;
    A.n = n;
    A.k = nidx;
    A.indexes = indexes;
    A.mol = moleculeIndex;
    A.elem = elementIndex;
    A.nel = numberOfElements;
    A.lower = lowerLimit;
    A.upper = upperLimit;
    A.flags = flags;
    A.allAtoms = allAtoms;
    // column c of the (n,k) array belongs to indexes[c]; the kernel indexes it with stride k
    (void)k;
    return ad_run<AD_DISTS>(A, dintra, dinter, nintra, ninter, as_stream(stream));
}

extern "C" int frmc_pair_elements_stats(const int32_t *elementIndex, const int32_t *moleculeIndex, int32_t n,
                                        int32_t numberOfElements, int32_t *nintra, int32_t *ninter, void *stream) {
    FRMC_REQUIRE(numberOfElements >= 1 && n >= 0, "pair_elements_stats: bad sizes");
    const size_t lds = sizeof(int32_t) * (2 * (size_t)numberOfElements * numberOfElements + 2 * kPeChunk);
    FRMC_REQUIRE(lds <= 64 * 1024, "pair_elements_stats: %d element types is too many", numberOfElements);
    hipStream_t st = as_stream(stream);
    const size_t nn = (size_t)numberOfElements * numberOfElements;
    FRMC_CHECK_HIP(hipMemsetAsync(nintra, 0, sizeof(int32_t) * nn, st));
    FRMC_CHECK_HIP(hipMemsetAsync(ninter, 0, sizeof(int32_t) * nn, st));
    if (n < 2) return FRMC_OK;
    const dim3 grid(blocks_for(n, 256), (n + kPeChunk - 1) / kPeChunk);
    hipLaunchKernelGGL(k_pair_elements_stats, grid, dim3(256), lds, st, elementIndex, moleculeIndex, n,
                       numberOfElements, nintra, ninter);
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}
