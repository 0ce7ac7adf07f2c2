// histograms.hip -- pairs_histograms entry points on gfx950.
//
// Counting is done in int32 (every reference increment is +1.0f on an
// integer-valued float32 bin, exact below 2^24) and converted to the
// reference's float32 arrays at the end.  Three kernels:
//
//  k_multi_coords  "atoms in a list vs all atoms": fused distance + bin, one
//                  thread per (other atom), the list atoms uniform per block;
//                  in-range pairs go straight to global int32 atomics (a
//                  move touches O(g*N*fraction_in_range) bins).
//  k_full_pairs    all unordered pairs i<j: triangular tile schedule, the
//                  row atom in registers, the column tile staged in LDS
//                  (float4 x,y,z,mol + elem), per-workgroup LDS-private bins
//                  (2*nel^2*H int32) merged once into global counts.
//  k_from_dists    binning of precomputed distance columns (the *_dists API).
//
// Slot convention (pairs_histograms.pyx:64-68): bin (elem(atom), elem(other),
// (int)((d-min)/bin)) of hintra if same molecule else hinter; flat C-order
// index, so an index equal to H aliases the next slot exactly like the
// reference; flat indices past the array are dropped and counted.
#include "frmc_device.h"
#include "frmc_host.h"

namespace frmc {

// LDS-private histogram limit per workgroup (gfx950 has 160 KB per CU)
static constexpr size_t kMaxLdsHistBytes = 96 * 1024;


__device__ __forceinline__ void load_basis_dev(const float *__restrict__ basis, Basis &B, bool &ortho) {
#pragma unroll
    for (int k = 0; k < 9; k++) B.b[k] = basis[k];
    ortho = (B.b[1] == 0.0f && B.b[2] == 0.0f && B.b[3] == 0.0f && B.b[5] == 0.0f && B.b[6] == 0.0f &&
             B.b[7] == 0.0f);
}

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

// --------------------------------------------------------------- zero/convert

__global__ void k_fill_i32(int32_t *__restrict__ p, int64_t n, int32_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// counts[0..size) -> hintra, counts[size..2size) -> hinter; add or overwrite
__global__ void k_counts_to_float(const int32_t *__restrict__ counts, int64_t size, float *__restrict__ hintra,
                                  float *__restrict__ hinter, int add) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < size; i += stride) {
        float a = (float)counts[i], b = (float)counts[size + i];
        if (add) {
            hintra[i] += a;
            hinter[i] += b;
        } else {
            hintra[i] = a;
            hinter[i] = b;
        }
    }
}

// counts layout: [2][nel][nel][H] int32 (intra block then inter block);
// counts[2*size] is the overflow counter.
__device__ __forceinline__ void count_pair(int32_t *__restrict__ counts, const HistArgs &A, float s, int32_t ea,
                                           int32_t eb, bool intra) {
    int32_t b = bin_of(s, A.g);
    int64_t flat = ((int64_t)ea * A.nel + eb) * A.g.H + b;
    if (flat < 0 || flat >= A.size) {
        atomicAdd(&counts[2 * A.size], 1);
        return;
    }
    atomicAdd(&counts[(intra ? 0 : A.size) + flat], 1);
}

// ------------------------------------------------------------ multi coords

// grid.x over other atoms i, grid.y over chunks of the index list.
template <bool PBC, bool ORTHO>
__global__ void __launch_bounds__(256) k_multi_coords(const int32_t *__restrict__ indexes, int32_t nidx,
                                                      int32_t per_block, const float *__restrict__ coords,
                                                      int32_t n, const float *__restrict__ basis,
                                                      const int32_t *__restrict__ mol,
                                                      const int32_t *__restrict__ elem, HistArgs A, int allAtoms,
                                                      int32_t *__restrict__ counts) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Basis B;
    bool ortho = false;
    if (PBC) load_basis_dev(basis, B, ortho);
    const float cx = coords[3 * (int64_t)i], cy = coords[3 * (int64_t)i + 1], cz = coords[3 * (int64_t)i + 2];
    int32_t mi = -1, ei = -1;
    const int32_t c0 = blockIdx.y * per_block;
    const int32_t c1 = min(nidx, c0 + per_block);
    for (int32_t c = c0; c < c1; c++) {
        const int32_t a = indexes[c];
        if (a < 0 || a >= n) continue;         // guarded; the host shim rejects such lists
        if (i == a) continue;                  // pairs_histograms.pyx:54
        if (!allAtoms && i < a) continue;      // start index = atomIndex
        const float ax = coords[3 * (int64_t)a], ay = coords[3 * (int64_t)a + 1], az = coords[3 * (int64_t)a + 2];
        float s = pair_s<PBC, ORTHO>(ax, ay, az, cx, cy, cz, B);
        if (!in_range(s, A.g)) continue;
        if (mi < 0) {
            mi = mol[i];
            ei = elem[i];
        }
        count_pair(counts, A, s, elem[a], ei, mol[a] == mi);
    }
}

// ------------------------------------------------------------- full pairs

// Work item = (row block bi, chunk of column blocks).  Row blocks are TILE
// atoms; column blocks are TILE atoms; only column blocks bj >= bi are used.
struct FullSched {
    int32_t nb;      // number of TILE blocks
    int32_t chunk;   // column blocks per work item
    int64_t items;   // total work items
};

template <int TILE, bool PBC, bool ORTHO, bool LDS_HIST>
__global__ void __launch_bounds__(TILE) k_full_pairs(const float *__restrict__ coords, int32_t n,
                                                     const float *__restrict__ basis,
                                                     const int32_t *__restrict__ mol,
                                                     const int32_t *__restrict__ elem, HistArgs A, FullSched S,
                                                     const int64_t *__restrict__ item_row_start,
                                                     int32_t *__restrict__ counts) {
    extern __shared__ int32_t lds_hist[];  // 2*size ints when LDS_HIST
    __shared__ float4 tile_xyzm[TILE];
    __shared__ int32_t tile_e[TILE];
    __shared__ int32_t lds_overflow;

    Basis B;
    bool ortho = false;
    if (PBC) load_basis_dev(basis, B, ortho);
    const int tid = threadIdx.x;
    if (LDS_HIST) {
        for (int64_t s = tid; s < 2 * A.size; s += TILE) lds_hist[s] = 0;
    }
    if (tid == 0) lds_overflow = 0;
    __syncthreads();

    for (int64_t item = blockIdx.x; item < S.items; item += gridDim.x) {
        // find the row block of this item: item_row_start[bi] <= item < item_row_start[bi+1]
        int32_t lo = 0, hi = S.nb;  // invariant: start[lo] <= item < start[hi]
        while (hi - lo > 1) {
            int32_t mid = (lo + hi) >> 1;
            if (item_row_start[mid] <= item) lo = mid; else hi = mid;
        }
        const int32_t bi = lo;
        const int32_t bj0 = bi + (int32_t)(item - item_row_start[bi]) * S.chunk;
        const int32_t bj1 = min(S.nb, bj0 + S.chunk);

        const int32_t i = bi * TILE + tid;
        const bool row_ok = i < n;
        float ix = 0.f, iy = 0.f, iz = 0.f;
        int32_t im = 0, ie = 0;
        if (row_ok) {
            ix = coords[3 * (int64_t)i];
            iy = coords[3 * (int64_t)i + 1];
            iz = coords[3 * (int64_t)i + 2];
            im = mol[i];
            ie = elem[i];
        }
        for (int32_t bj = bj0; bj < bj1; bj++) {
            __syncthreads();  // previous tile consumed
            {
                const int32_t j = bj * TILE + tid;
                if (j < n) {
                    tile_xyzm[tid] = make_float4(coords[3 * (int64_t)j], coords[3 * (int64_t)j + 1],
                                                 coords[3 * (int64_t)j + 2], __int_as_float(mol[j]));
                    tile_e[tid] = elem[j];
                }
            }
            __syncthreads();
            if (!row_ok) continue;
            const int32_t jbase = bj * TILE;
            const int32_t jn = min(TILE, n - jbase);
            // diagonal tile: only j > i, i.e. t > tid
            int32_t t0 = (bj == bi) ? tid + 1 : 0;
#pragma unroll 4
            for (int32_t t = t0; t < jn; t++) {
                const float4 q = tile_xyzm[t];
                // reference: atomIndex is the earlier atom i, diff = c_i - c_j
                float s = pair_s<PBC, ORTHO>(ix, iy, iz, q.x, q.y, q.z, B);
                if (!in_range(s, A.g)) continue;
                int32_t b = bin_of(s, A.g);
                int64_t flat = ((int64_t)ie * A.nel + tile_e[t]) * A.g.H + b;
                bool intra = __float_as_int(q.w) == im;
                if (flat < 0 || flat >= A.size) {
                    atomicAdd(&lds_overflow, 1);
                } else if (LDS_HIST) {
                    atomicAdd(&lds_hist[(intra ? 0 : A.size) + flat], 1);
                } else {
                    atomicAdd(&counts[(intra ? 0 : A.size) + flat], 1);
                }
            }
        }
    }
    __syncthreads();
    if (LDS_HIST) {
        for (int64_t s = tid; s < 2 * A.size; s += TILE) {
            int32_t v = lds_hist[s];
            if (v) atomicAdd(&counts[s], v);
        }
    }
    if (tid == 0 && lds_overflow) atomicAdd(&counts[2 * A.size], lds_overflow);
}

// --------------------------------------------------------------- from dists

__global__ void __launch_bounds__(256) k_from_dists(const int32_t *__restrict__ indexes, int32_t atomIndex,
                                                    const float *__restrict__ dist, int32_t n, int32_t k,
                                                    const int32_t *__restrict__ mol,
                                                    const int32_t *__restrict__ elem, HistArgs A, int allAtoms,
                                                    int32_t *__restrict__ counts) {
    const int32_t c = blockIdx.y;
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t a = indexes ? indexes[c] : atomIndex;
    if (a < 0 || a >= n) return;  // guarded; the host shim rejects such lists
    if (i == a) return;
    if (!allAtoms && i < a) return;
    float d = dist[(int64_t)i * k + c];
    // pairs_histograms.pyx:58-64 on a given distance
    if (!(d >= A.g.minD) || d >= A.g.maxD) return;  // NaN dropped with d<min
    int32_t b = (int32_t)cr_div(d - A.g.minD, A.g.bin);
    int64_t flat = ((int64_t)elem[a] * A.nel + elem[i]) * A.g.H + b;
    if (flat < 0 || flat >= A.size) {
        atomicAdd(&counts[2 * A.size], 1);
        return;
    }
    atomicAdd(&counts[(mol[i] == mol[a] ? 0 : A.size) + flat], 1);
}

// ------------------------------------------------------------------- host

static int make_args(int32_t nel, int32_t H, float minD, float maxD, float bin, HistArgs &A) {
    FRMC_REQUIRE(nel >= 1 && H >= 1, "numberOfElements (%d) and histSize (%d) must be >= 1", nel, H);
    FRMC_REQUIRE((int64_t)nel * nel * H < (int64_t)1 << 30, "histogram too large (%d elements, %d bins)", nel, H);
    A.g = make_grid(minD, maxD, bin, H);
    A.nel = nel;
    A.size = (int64_t)nel * nel * H;
    return FRMC_OK;
}

// Scratch int32 counts [2*size + 1] from the stream-ordered pool.
struct Scratch {
    int32_t *p = nullptr;
    hipStream_t st = nullptr;
    ~Scratch() {
        if (p) (void)hipFreeAsync(p, st);
    }
};

static int alloc_counts(Scratch &sc, int64_t size, hipStream_t st) {
    sc.st = st;
    hipError_t e = hipMallocAsync((void **)&sc.p, sizeof(int32_t) * (2 * size + 1), st);
    if (e != hipSuccess) {
        set_error("hipMallocAsync of %lld histogram counts failed: %s", (long long)(2 * size + 1),
                  hipGetErrorString(e));
        return FRMC_ENOMEM;
    }
    k_fill_i32<<<blocks_for(2 * size + 1, 256) > 1024 ? 1024 : blocks_for(2 * size + 1, 256), 256, 0, st>>>(
        sc.p, 2 * size + 1, 0);
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

static int finish_counts(Scratch &sc, int64_t size, float *hintra, float *hinter, int add, hipStream_t st) {
    unsigned g = blocks_for(size, 256);
    if (g > 2048) g = 2048;
    k_counts_to_float<<<g, 256, 0, st>>>(sc.p, size, hintra, hinter, add);
    FRMC_CHECK_LAUNCH();
    int32_t ov = 0;
    FRMC_CHECK_HIP(hipMemcpyAsync(&ov, sc.p + 2 * size, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FRMC_CHECK_HIP(hipStreamSynchronize(st));
    set_overflow(ov);
    return FRMC_OK;
}

static int launch_multi(const int32_t *indexes, int32_t nidx, const float *coords, int32_t n, const float *basis,
                        int isPBC, const int32_t *mol, const int32_t *elem, const HistArgs &A, int allAtoms,
                        int32_t *counts, hipStream_t st, bool ortho_hint) {
    if (nidx == 0 || n == 0) return FRMC_OK;
    const int32_t per_block = 16;
    dim3 grid(blocks_for(n, 256), (unsigned)((nidx + per_block - 1) / per_block));
    FRMC_REQUIRE(grid.y <= 65535, "index list too long (%d)", nidx);
    if (isPBC) {
        if (ortho_hint)
            k_multi_coords<true, true><<<grid, 256, 0, st>>>(indexes, nidx, per_block, coords, n, basis, mol, elem,
                                                             A, allAtoms, counts);
        else
            k_multi_coords<true, false><<<grid, 256, 0, st>>>(indexes, nidx, per_block, coords, n, basis, mol,
                                                              elem, A, allAtoms, counts);
    } else {
        k_multi_coords<false, false><<<grid, 256, 0, st>>>(indexes, nidx, per_block, coords, n, basis, mol, elem, A,
                                                           allAtoms, counts);
    }
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

// The orthorhombic fast path needs the basis on the host; callers that only
// hold a device basis read it back once (36 bytes).
static int read_ortho(const float *basis, int isPBC, hipStream_t st, bool &ortho) {
    ortho = false;
    if (!isPBC) return FRMC_OK;
    float b[9];
    FRMC_CHECK_HIP(hipMemcpyAsync(b, basis, sizeof(b), hipMemcpyDeviceToHost, st));
    FRMC_CHECK_HIP(hipStreamSynchronize(st));
    ortho = b[1] == 0.f && b[2] == 0.f && b[3] == 0.f && b[5] == 0.f && b[6] == 0.f && b[7] == 0.f;
    return FRMC_OK;
}

template <int TILE>
static int launch_full_t(const float *coords, int32_t n, const float *basis, int isPBC, bool ortho,
                         const int32_t *mol, const int32_t *elem, const HistArgs &A, int32_t *counts,
                         hipStream_t st) {
    FullSched S;
    S.nb = (n + TILE - 1) / TILE;
    S.chunk = 8;
    // row-start prefix of work items (host-computed, uploaded)
    int64_t *h_start = new int64_t[S.nb + 1];
    int64_t acc = 0;
    for (int32_t bi = 0; bi < S.nb; bi++) {
        h_start[bi] = acc;
        acc += (S.nb - bi + S.chunk - 1) / S.chunk;
    }
    h_start[S.nb] = acc;
    S.items = acc;
    int64_t *d_start = nullptr;
    hipError_t e = hipMallocAsync((void **)&d_start, sizeof(int64_t) * (S.nb + 1), st);
    if (e != hipSuccess) {
        delete[] h_start;
        set_error("hipMallocAsync failed: %s", hipGetErrorString(e));
        return FRMC_ENOMEM;
    }
    e = hipMemcpyAsync(d_start, h_start, sizeof(int64_t) * (S.nb + 1), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);  // h_start is pageable and freed below
    delete[] h_start;
    if (e != hipSuccess) {
        (void)hipFreeAsync(d_start, st);
        set_error("upload of the tile schedule failed: %s", hipGetErrorString(e));
        return FRMC_EHIP;
    }
    const size_t lds_bytes = sizeof(int32_t) * 2 * A.size;
    const bool lds = lds_bytes <= kMaxLdsHistBytes;
    int dev = 0, ncu = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    // 160 KB of LDS per CU: as many resident blocks as the private bins allow
    int per_cu = lds ? (int)((150 * 1024) / (lds_bytes + 6 * 1024)) : 4;
    if (per_cu < 1) per_cu = 1;
    if (per_cu > 4) per_cu = 4;
    int64_t blocks = (int64_t)ncu * per_cu;
    if (blocks > S.items) blocks = S.items;
    if (blocks < 1) blocks = 1;
    size_t shm = lds ? lds_bytes : 0;
#define FRMC_FULL_LAUNCH(P, O, L)                                                                               \
    do {                                                                                                        \
        if (shm > 48 * 1024)                                                                                    \
            (void)hipFuncSetAttribute((const void *)k_full_pairs<TILE, P, O, L>,                                \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);                    \
        k_full_pairs<TILE, P, O, L><<<(unsigned)blocks, TILE, shm, st>>>(coords, n, basis, mol, elem, A, S,   \
                                                                         d_start, counts);                      \
    } while (0)
    if (isPBC) {
        if (ortho) { if (lds) FRMC_FULL_LAUNCH(true, true, true); else FRMC_FULL_LAUNCH(true, true, false); }
        else       { if (lds) FRMC_FULL_LAUNCH(true, false, true); else FRMC_FULL_LAUNCH(true, false, false); }
    } else {
        if (lds) FRMC_FULL_LAUNCH(false, false, true); else FRMC_FULL_LAUNCH(false, false, false);
    }
#undef FRMC_FULL_LAUNCH
    hipError_t le = hipGetLastError();
    (void)hipFreeAsync(d_start, st);
    if (le != hipSuccess) {
        set_error("k_full_pairs launch failed: %s", hipGetErrorString(le));
        return FRMC_EHIP;
    }
    return FRMC_OK;
}

int launch_full(const float *coords, int32_t n, const float *basis, int isPBC, bool ortho, const int32_t *mol,
                const int32_t *elem, const HistArgs &A, int32_t *counts, hipStream_t st) {
    if (n < 2) return FRMC_OK;
    hipEvent_t a = nullptr, b = nullptr;
    const bool timing = timing_enabled();
    if (timing) {
        FRMC_CHECK_HIP(hipEventCreate(&a));
        FRMC_CHECK_HIP(hipEventCreate(&b));
        FRMC_CHECK_HIP(hipEventRecord(a, st));
    }
    int rc = launch_full_t<256>(coords, n, basis, isPBC, ortho, mol, elem, A, counts, st);
    if (timing) {
        float ms = 0.f;
        hipError_t e = hipEventRecord(b, st);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        if (e != hipSuccess && rc == FRMC_OK) {
            set_error("kernel timing failed: %s", hipGetErrorString(e));
            rc = FRMC_EHIP;
        }
        set_last_kernel_ms(ms);
    }
    return rc;
}

}  // namespace frmc

using namespace frmc;

extern "C" {

int frmc_single_pairs_histograms(int32_t atomIndex, const float *distances, int32_t n, const int32_t *mol,
                                 const int32_t *elem, float *hintra, float *hinter, int32_t nel, int32_t H,
                                 float minD, float maxD, float bin, int allAtoms, void *stream) {
    HistArgs A;
    int rc = make_args(nel, H, minD, maxD, bin, A);
    if (rc) return rc;
    FRMC_REQUIRE(n >= 0, "negative size");
    FRMC_REQUIRE(atomIndex >= 0 && atomIndex < n, "atomIndex %d out of range [0,%d)", atomIndex, n);
    FRMC_REQUIRE(distances && mol && elem && hintra && hinter, "NULL pointer argument");
    hipStream_t st = as_stream(stream);
    Scratch sc;
    if ((rc = alloc_counts(sc, A.size, st))) return rc;
    dim3 grid(blocks_for(n, 256), 1);
    k_from_dists<<<grid, 256, 0, st>>>(nullptr, atomIndex, distances, n, 1, mol, elem, A, allAtoms, sc.p);
    FRMC_CHECK_LAUNCH();
    return finish_counts(sc, A.size, hintra, hinter, /*add=*/1, st);
}

int frmc_multiple_pairs_histograms_coords(const int32_t *indexes, int32_t nidx, const float *coords, int32_t n,
                                          const float *basis, int isPBC, const int32_t *mol, const int32_t *elem,
                                          int32_t nel, float minD, float maxD, float bin, int32_t H, int allAtoms,
                                          float *hintra, float *hinter, void *stream) {
    HistArgs A;
    int rc = make_args(nel, H, minD, maxD, bin, A);
    if (rc) return rc;
    FRMC_REQUIRE(n >= 0 && nidx >= 0, "negative size");
    FRMC_REQUIRE(hintra && hinter, "NULL histogram pointer");
    FRMC_REQUIRE(n == 0 || (coords && mol && elem), "NULL pointer argument");
    FRMC_REQUIRE(nidx == 0 || indexes, "indexes is NULL");
    FRMC_REQUIRE(!isPBC || basis, "basis is NULL");
    hipStream_t st = as_stream(stream);
    bool ortho = false;
    if ((rc = read_ortho(basis, isPBC, st, ortho))) return rc;
    Scratch sc;
    if ((rc = alloc_counts(sc, A.size, st))) return rc;
    if ((rc = launch_multi(indexes, nidx, coords, n, basis, isPBC, mol, elem, A, allAtoms, sc.p, st, ortho)))
        return rc;
    return finish_counts(sc, A.size, hintra, hinter, /*add=*/0, st);
}

int frmc_multiple_pairs_histograms_dists(const int32_t *indexes, int32_t nidx, const float *distances, int32_t n,
                                         int32_t k, const int32_t *mol, const int32_t *elem, int32_t nel,
                                         float minD, float maxD, float bin, int32_t H, int allAtoms, float *hintra,
                                         float *hinter, void *stream) {
    HistArgs A;
    int rc = make_args(nel, H, minD, maxD, bin, A);
    if (rc) return rc;
    FRMC_REQUIRE(n >= 0 && nidx >= 0 && k >= nidx, "bad sizes (n=%d, nidx=%d, k=%d)", n, nidx, k);
    FRMC_REQUIRE(hintra && hinter, "NULL histogram pointer");
    FRMC_REQUIRE(nidx <= 65535, "index list too long (%d)", nidx);
    hipStream_t st = as_stream(stream);
    Scratch sc;
    if ((rc = alloc_counts(sc, A.size, st))) return rc;
    if (nidx > 0 && n > 0) {
        FRMC_REQUIRE(indexes && distances && mol && elem, "NULL pointer argument");
        dim3 grid(blocks_for(n, 256), (unsigned)nidx);
        k_from_dists<<<grid, 256, 0, st>>>(indexes, 0, distances, n, k, mol, elem, A, allAtoms, sc.p);
        FRMC_CHECK_LAUNCH();
    }
    return finish_counts(sc, A.size, hintra, hinter, /*add=*/0, st);
}

int frmc_full_pairs_histograms_coords(const float *coords, int32_t n, const float *basis, int isPBC,
                                      const int32_t *mol, const int32_t *elem, int32_t nel, float minD, float maxD,
                                      float bin, int32_t H, float *hintra, float *hinter, void *stream) {
    HistArgs A;
    int rc = make_args(nel, H, minD, maxD, bin, A);
    if (rc) return rc;
    FRMC_REQUIRE(n >= 0, "negative size");
    FRMC_REQUIRE(hintra && hinter, "NULL histogram pointer");
    FRMC_REQUIRE(n == 0 || (coords && mol && elem), "NULL pointer argument");
    FRMC_REQUIRE(!isPBC || basis, "basis is NULL");
    hipStream_t st = as_stream(stream);
    bool ortho = false;
    if ((rc = read_ortho(basis, isPBC, st, ortho))) return rc;
    Scratch sc;
    if ((rc = alloc_counts(sc, A.size, st))) return rc;
    if ((rc = launch_full(coords, n, basis, isPBC, ortho, mol, elem, A, sc.p, st))) return rc;
    return finish_counts(sc, A.size, hintra, hinter, /*add=*/0, st);
}

int frmc_full_pairs_histograms_dists(const float *distances, int32_t n, int32_t k, const int32_t *mol,
                                     const int32_t *elem, int32_t nel, float minD, float maxD, float bin, int32_t H,
                                     float *hintra, float *hinter, void *stream) {
    // pairs_histograms.pyx:377-383: indexes = arange(k), allAtoms=False
    HistArgs A;
    int rc = make_args(nel, H, minD, maxD, bin, A);
    if (rc) return rc;
    FRMC_REQUIRE(n >= 0 && k >= 0, "negative size");
    FRMC_REQUIRE(k <= n || n == 0, "distances has more columns (%d) than atoms (%d)", k, n);
    FRMC_REQUIRE(k <= 65535, "too many columns (%d)", k);
    FRMC_REQUIRE(hintra && hinter, "NULL histogram pointer");
    hipStream_t st = as_stream(stream);
    Scratch sc;
    if ((rc = alloc_counts(sc, A.size, st))) return rc;
    if (n > 0 && k > 0) {
        FRMC_REQUIRE(distances && mol && elem, "NULL pointer argument");
        // column c is atom c: use an identity index list through atomIndex per column
        int32_t *idx = nullptr;
        hipError_t e = hipMallocAsync((void **)&idx, sizeof(int32_t) * k, st);
        if (e != hipSuccess) {
            set_error("hipMallocAsync failed: %s", hipGetErrorString(e));
            return FRMC_ENOMEM;
        }
        int32_t *h = new int32_t[k];
        for (int32_t c = 0; c < k; c++) h[c] = c;
        e = hipMemcpyAsync(idx, h, sizeof(int32_t) * k, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        delete[] h;
        if (e != hipSuccess) {
            (void)hipFreeAsync(idx, st);
            set_error("index upload failed: %s", hipGetErrorString(e));
            return FRMC_EHIP;
        }
        dim3 grid(blocks_for(n, 256), (unsigned)k);
        k_from_dists<<<grid, 256, 0, st>>>(idx, 0, distances, n, k, mol, elem, A, 0, sc.p);
        hipError_t le = hipGetLastError();
        (void)hipFreeAsync(idx, st);
        if (le != hipSuccess) {
            set_error("k_from_dists launch failed: %s", hipGetErrorString(le));
            return FRMC_EHIP;
        }
    }
    return finish_counts(sc, A.size, hintra, hinter, /*add=*/0, st);
}

int frmc_full_pairs_counts(const float *coords, int32_t n, const float *basis, int isPBC, const int32_t *mol,
                           const int32_t *elem, int32_t nel, float minD, float maxD, float bin, int32_t H,
                           int32_t *counts, void *stream) {
    HistArgs A;
    int rc = make_args(nel, H, minD, maxD, bin, A);
    if (rc) return rc;
    FRMC_REQUIRE(n >= 0 && counts, "bad arguments");
    FRMC_REQUIRE(n == 0 || (coords && mol && elem), "NULL pointer argument");
    FRMC_REQUIRE(!isPBC || basis, "basis is NULL");
    hipStream_t st = as_stream(stream);
    bool ortho = false;
    if ((rc = read_ortho(basis, isPBC, st, ortho))) return rc;
    k_fill_i32<<<1024, 256, 0, st>>>(counts, 2 * A.size + 1, 0);
    FRMC_CHECK_LAUNCH();
    return launch_full(coords, n, basis, isPBC, ortho, mol, elem, A, counts, st);
}

}  // extern "C"
