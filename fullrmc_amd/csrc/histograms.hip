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
#include <math.h>
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

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

// Work item = (row block bi of R*NT atoms, column tile c of NT atoms), NT the
// workgroup size, with c >= R*bi: the upper triangle of tiles, diagonal tiles
// included (there only j > i is counted).  Items are numbered row block by
// row block, so row block bi starts at item bi*nct - R*bi*(bi-1)/2; a
// workgroup takes a contiguous range of items and keeps its row atoms in
// registers while bi is unchanged.

struct FullSched {
    int32_t R;          // row atoms per thread
    int32_t nrb, nct;   // row blocks, column tiles
    int64_t items;      // work items
    int64_t per_block;  // items per workgroup
};

__host__ __device__ __forceinline__ int64_t row_block_start(int64_t bi, int64_t nct, int64_t R) {
    return bi * nct - R * bi * (bi - 1) / 2;
}

// Float64 edge re-check (EDGE mode): the histogram of the binary64 recompute
// (every pair's distance evaluated in double from the float32 coordinates,
// range and bin in double).  The float32 distance d32 of a pair is within
// eps = 8u*(sum_m (|dx_m|+1)*rowsum_m + d32) of the double one (u = 2^-24;
// error analysis of the float32 difference, the exact minimum-image
// subtraction, the 3-term basis product, the squared norm and the rounded
// sqrt, with a factor > 2 to spare); a pair whose d32 lies farther than eps
// from every bin edge and range end, and whose fractional components are not
// within a rounding tie of +-1/2, takes its bin from d32; every other pair is
// recomputed in double.
struct Edge64 {
    double mn, mx, bn;      // the grid in double
    float s_lo, s_hi;       // candidate window of float32 squared norms
    float rowsum[3];        // sum_k |B[m][k]| per basis row (PBC), 1 otherwise
};

__device__ __forceinline__ double round64(double x) { return x > 0.0 ? floor(x + 0.5) : ceil(x - 0.5); }

template <bool PBC>
__device__ __forceinline__ double pair_d64(float ax, float ay, float az, float cx, float cy, float cz,
                                           const Basis &B) {
    const double dx = (double)ax - (double)cx, dy = (double)ay - (double)cy, dz = (double)az - (double)cz;
    double rx = dx, ry = dy, rz = dz;
    if (PBC) {
        const double bx = dx - round64(dx), by = dy - round64(dy), bz = dz - round64(dz);
        rx = (bx * (double)B.b[0] + by * (double)B.b[3]) + bz * (double)B.b[6];
        ry = (bx * (double)B.b[1] + by * (double)B.b[4]) + bz * (double)B.b[7];
        rz = (bx * (double)B.b[2] + by * (double)B.b[5]) + bz * (double)B.b[8];
    }
    return __builtin_sqrt((rx * rx + ry * ry) + rz * rz);
}

// Bin of a candidate pair in the binary64 recompute, or -1 out of range;
// *rechecked is set when the pair needed the double evaluation.
template <bool PBC>
__device__ __forceinline__ int64_t edge_bin(float s, float ax, float ay, float az, float cx, float cy, float cz,
                                            const Basis &B, const Edge64 &E, bool &rechecked) {
    const float dx = ax - cx, dy = ay - cy, dz = az - cz;
    const float d32 = cr_sqrt(s);
    const float bound = (fabsf(dx) + 1.0f) * E.rowsum[0] + (fabsf(dy) + 1.0f) * E.rowsum[1] +
                        (fabsf(dz) + 1.0f) * E.rowsum[2] + d32;
    const double eps = 8.0 * 5.9604644775390625e-08 * 1.001 * (double)bound;
    bool tie = false;
    if (PBC) {
        // fractional differences near +-1/2: float and double may pick different images
        tie = fabsf(0.5f - fabsf(min_image(dx))) <= 4.8e-7f * (fabsf(dx) + 1.0f) ||
              fabsf(0.5f - fabsf(min_image(dy))) <= 4.8e-7f * (fabsf(dy) + 1.0f) ||
              fabsf(0.5f - fabsf(min_image(dz))) <= 4.8e-7f * (fabsf(dz) + 1.0f);
    }
    const double t = ((double)d32 - E.mn) / E.bn;
    const double kf = floor(t);
    const double margin = fmin(t - kf, kf + 1.0 - t) * E.bn;
    rechecked = tie || margin <= eps || fabs((double)d32 - E.mn) <= eps || fabs((double)d32 - E.mx) <= eps;
    if (rechecked) {
        const double d = pair_d64<PBC>(ax, ay, az, cx, cy, cz, B);
        if (d < E.mn || d >= E.mx) return -1;
        return (int64_t)((d - E.mn) / E.bn);
    }
    if ((double)d32 < E.mn || (double)d32 >= E.mx) return -1;
    return (int64_t)kf;
}

template <int NT, int R, bool PBC, bool ORTHO, bool LDS_HIST, bool EDGE>
__global__ void __launch_bounds__(NT) k_full_pairs(const float *__restrict__ coords, int32_t n, Basis B,
                                                    const int32_t *__restrict__ mol,
                                                    const int32_t *__restrict__ elem, HistArgs A, FullSched S,
                                                    Edge64 E, int32_t *__restrict__ partial,
                                                    int32_t *__restrict__ counts) {
    extern __shared__ int32_t lds_hist[];  // 2*size ints when LDS_HIST
    __shared__ float4 tile_xyzm[NT];
    __shared__ int32_t tile_e[NT];
    __shared__ int32_t lds_overflow;
    __shared__ unsigned long long lds_rechecked;

    const int tid = threadIdx.x;
    const int32_t size = (int32_t)A.size;  // LDS_HIST: 2*size ints fit in LDS
    if (LDS_HIST)
        for (int32_t s = tid; s < 2 * size; s += NT) lds_hist[s] = 0;
    if (tid == 0) {
        lds_overflow = 0;
        lds_rechecked = 0;
    }

    const int64_t it0 = (int64_t)blockIdx.x * S.per_block;
    const int64_t it1 = min(S.items, it0 + S.per_block);
    // row block of the first item: largest bi with start(bi) <= it0
    int32_t lo = 0, hi = S.nrb;
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (row_block_start(mid, S.nct, R) <= it0) lo = mid; else hi = mid;
    }
    int32_t bi = lo, cur_bi = -1;
    float rx[R], ry[R], rz[R];
    int32_t rm[R], re[R], ri[R];
    int32_t lost = 0;
    uint32_t nre = 0;
    for (int64_t item = it0; item < it1; item++) {
        while (bi + 1 < S.nrb && item >= row_block_start(bi + 1, S.nct, R)) bi++;
        const int32_t c = R * bi + (int32_t)(item - row_block_start(bi, S.nct, R));
        if (bi != cur_bi) {
            cur_bi = bi;
#pragma unroll
            for (int r = 0; r < R; r++) {
                ri[r] = (bi * R + r) * NT + tid;
                const int32_t i = min(ri[r], n - 1);  // clamped; rows >= n never count
                rx[r] = coords[3 * (int64_t)i];
                ry[r] = coords[3 * (int64_t)i + 1];
                rz[r] = coords[3 * (int64_t)i + 2];
                rm[r] = mol[i];
                re[r] = elem[i];
            }
        }
        __syncthreads();  // previous tile consumed (and LDS bins cleared)
        {
            const int32_t j = min(c * NT + tid, n - 1);
            tile_xyzm[tid] = make_float4(coords[3 * (int64_t)j], coords[3 * (int64_t)j + 1],
                                         coords[3 * (int64_t)j + 2], __int_as_float(mol[j]));
            tile_e[tid] = elem[j];
        }
        __syncthreads();
        const int32_t jbase = c * NT;
        const int32_t jn = min(NT, n - jbase);
        const bool diag = c < R * (bi + 1);  // tile overlaps the row block
#pragma unroll 2
        for (int32_t t = 0; t < jn; t++) {
            const float4 q = tile_xyzm[t];
            const int32_t ej = tile_e[t];
            const int32_t j = jbase + t;
#pragma unroll
            for (int r = 0; r < R; r++) {
                // the earlier atom i is the reference's atomIndex: diff = c_i - c_j
                const float s = pair_s<PBC, ORTHO>(rx[r], ry[r], rz[r], q.x, q.y, q.z, B);
                const bool cand = EDGE ? (s >= E.s_lo && s < E.s_hi) : in_range(s, A.g);
                const bool ok = cand && ri[r] < n && (!diag || j > ri[r]);
                if (ok) {
                    int64_t b;
                    if (EDGE) {
                        bool re64 = false;
                        b = edge_bin<PBC>(s, rx[r], ry[r], rz[r], q.x, q.y, q.z, B, E, re64);
                        nre += re64;
                        if (b < 0) continue;
                    } else {
                        b = bin_of(s, A.g);
                    }
                    const int64_t flat = ((int64_t)re[r] * A.nel + ej) * A.g.H + b;
                    const bool intra = __float_as_int(q.w) == rm[r];
                    if (flat >= A.size) {
                        lost++;
                    } else if (LDS_HIST) {
                        atomicAdd(&lds_hist[(intra ? 0 : size) + (int32_t)flat], 1);
                    } else {
                        atomicAdd(&counts[(intra ? 0 : A.size) + flat], 1);
                    }
                }
            }
        }
    }
    if (lost) atomicAdd(&lds_overflow, lost);
    if (EDGE && nre) atomicAdd(&lds_rechecked, (unsigned long long)nre);
    __syncthreads();
    if (LDS_HIST) {
        int32_t *dst = partial + (int64_t)blockIdx.x * 2 * size;
        for (int32_t s = tid; s < 2 * size; s += NT) dst[s] = lds_hist[s];
    }
    if (tid == 0 && lds_overflow) atomicAdd(&counts[2 * A.size], lds_overflow);
    // EDGE: counts[2*size+1 .. +2] hold the number of double re-checks (64-bit, low word first)
    if (EDGE && tid == 0 && lds_rechecked)
        atomicAdd(reinterpret_cast<unsigned long long *>(&counts[2 * A.size + 2]), lds_rechecked);
}

// ------------------------------------------------ cell-list full histogram
//
// For periodic cells much larger than the histogram range, most pairs are
// out of range.  The atoms are sorted into a grid of m0 x m1 x m2 fractional
// cells, each at least the range (plus a margin) across: an atom's in-range
// partners then lie in its own or the 26 neighbouring cells (the margin
// covers the float32 rounding of the difference, image shift and norm).  A
// work item is (cell c, one of 14 offsets: itself and 13 of the 26
// neighbours, one of each opposite pair), so every unordered cell pair is
// visited once; pairs whose cells are not neighbours are never evaluated and
// count nowhere in either arithmetic.  Each pair keeps the reference's slot
// [element(earlier index)][element(later index)] and intra/inter by molecule;
// its distance is the same float32 (or binary64) value in either order.
struct CellGrid {
    int32_t m0, m1, m2;
    int64_t ncell;
    int64_t items;      // 14 per cell
    int64_t per_block;  // items per workgroup
};

__device__ __forceinline__ int32_t cell_of(float x, int32_t m) {
    const float f = x - floorf(x);  // exact: the fraction of a float
    const int32_t c = (int32_t)(f * (float)m);
    return c < 0 ? 0 : (c >= m ? m - 1 : c);
}

__global__ void __launch_bounds__(256) k_cell_count(const float *__restrict__ coords, int32_t n, CellGrid G,
                                                    int32_t *__restrict__ cid, int32_t *__restrict__ cnt) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t cx = cell_of(coords[3 * (int64_t)i], G.m0), cy = cell_of(coords[3 * (int64_t)i + 1], G.m1),
                  cz = cell_of(coords[3 * (int64_t)i + 2], G.m2);
    const int32_t c = (cx * G.m1 + cy) * G.m2 + cz;
    cid[i] = c;
    atomicAdd(&cnt[c], 1);
}

// exclusive scan of ncell counts into start[0..ncell] (one block); fill <- start
__global__ void __launch_bounds__(1024) k_cell_scan(const int32_t *__restrict__ cnt, int64_t ncell,
                                                    int32_t *__restrict__ start, int32_t *__restrict__ fill) {
    __shared__ int32_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (ncell + 1023) / 1024;
    const int64_t c0 = t * per, c1 = min(ncell, c0 + per);
    int32_t sum = 0;
    for (int64_t c = c0; c < c1; c++) sum += cnt[c];
    part[t] = sum;
    __syncthreads();
    if (t == 0) {
        int32_t acc = 0;
        for (int q = 0; q < 1024; q++) {
            const int32_t v = part[q];
            part[q] = acc;
            acc += v;
        }
        start[ncell] = acc;
    }
    __syncthreads();
    int32_t acc = part[t];
    for (int64_t c = c0; c < c1; c++) {
        start[c] = acc;
        fill[c] = acc;
        acc += cnt[c];
    }
}

// cell-sorted copies: (x, y, z, molecule bits), element, original index
__global__ void __launch_bounds__(256) k_cell_scatter(const float *__restrict__ coords, int32_t n,
                                                      const int32_t *__restrict__ mol,
                                                      const int32_t *__restrict__ elem,
                                                      const int32_t *__restrict__ cid, int32_t *__restrict__ fill,
                                                      float4 *__restrict__ sx, int32_t *__restrict__ se,
                                                      int32_t *__restrict__ sid) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t pos = atomicAdd(&fill[cid[i]], 1);
    sx[pos] = make_float4(coords[3 * (int64_t)i], coords[3 * (int64_t)i + 1], coords[3 * (int64_t)i + 2],
                          __int_as_float(mol[i]));
    se[pos] = elem[i];
    sid[pos] = i;
}

// offset o of the 14 visited per cell: 0 the cell itself, 1..13 the
// neighbours (dx, dy, dz) with dx > 0, or dx == 0 and dy > 0, or dx == dy == 0
// and dz > 0
__device__ __forceinline__ void cell_offset(int32_t o, int32_t &dx, int32_t &dy, int32_t &dz) {
    if (o == 0) {
        dx = dy = dz = 0;
        return;
    }
    if (o <= 9) {  // dx = 1: the 9 (dy, dz)
        dx = 1;
        dy = (o - 1) / 3 - 1;
        dz = (o - 1) % 3 - 1;
    } else if (o <= 12) {  // dx = 0, dy = 1
        dx = 0;
        dy = 1;
        dz = o - 11;
    } else {  // dx = dy = 0, dz = 1
        dx = 0;
        dy = 0;
        dz = 1;
    }
}

// NT threads, up to R rows per thread per pass (a cell of more than R*NT
// atoms takes several passes); LDS_HIST: workgroup-private bins as in
// k_full_pairs, written to this workgroup's partial at the end.
template <int NT, int R, bool ORTHO, bool LDS_HIST, bool EDGE>
__global__ void __launch_bounds__(NT) k_full_cells(const float4 *__restrict__ sx, const int32_t *__restrict__ se,
                                                    const int32_t *__restrict__ sid,
                                                    const int32_t *__restrict__ cstart, CellGrid G, Basis B,
                                                    HistArgs A, Edge64 E, int32_t *__restrict__ partial,
                                                    int32_t *__restrict__ counts) {
    extern __shared__ int32_t lds_hist[];
    __shared__ float4 tile_x[NT];
    __shared__ int32_t tile_e[NT], tile_id[NT];
    __shared__ int32_t lds_overflow;
    __shared__ unsigned long long lds_rechecked;
    const int tid = threadIdx.x;
    const int32_t size = (int32_t)A.size;
    if (LDS_HIST)
        for (int32_t s = tid; s < 2 * size; s += NT) lds_hist[s] = 0;
    if (tid == 0) {
        lds_overflow = 0;
        lds_rechecked = 0;
    }
    const int64_t it0 = (int64_t)blockIdx.x * G.per_block;
    const int64_t it1 = min(G.items, it0 + G.per_block);
    int32_t lost = 0;
    uint32_t nre = 0;
    for (int64_t item = it0; item < it1; item++) {
        const int64_t c = item / 14;
        const int32_t o = (int32_t)(item % 14);
        const int32_t cx = (int32_t)(c / ((int64_t)G.m1 * G.m2)), cy = (int32_t)((c / G.m2) % G.m1),
                      cz = (int32_t)(c % G.m2);
        int32_t dx, dy, dz;
        cell_offset(o, dx, dy, dz);
        const int64_t c2 = ((int64_t)((cx + dx + G.m0) % G.m0) * G.m1 + (cy + dy + G.m1) % G.m1) * G.m2 +
                           (cz + dz + G.m2) % G.m2;
        const bool self = o == 0;
        const int32_t r0 = cstart[c], r1 = cstart[c + 1], q0 = cstart[c2], q1 = cstart[c2 + 1];
        for (int32_t rp = r0; rp < r1; rp += NT * R) {
            float rx[R], ry[R], rz[R];
            int32_t rm[R], re[R], ri[R], rpos[R];
#pragma unroll
            for (int k = 0; k < R; k++) {
                rpos[k] = rp + tid + NT * k;
                const int32_t pr = min(rpos[k], r1 - 1);  // clamped; rows past the cell never count
                const float4 v = sx[pr];
                rx[k] = v.x;
                ry[k] = v.y;
                rz[k] = v.z;
                rm[k] = __float_as_int(v.w);
                re[k] = se[pr];
                ri[k] = sid[pr];
            }
            for (int32_t tq = q0; tq < q1; tq += NT) {
                __syncthreads();  // previous tile consumed
                {
                    const int32_t j = min(tq + tid, q1 - 1);
                    tile_x[tid] = sx[j];
                    tile_e[tid] = se[j];
                    tile_id[tid] = sid[j];
                }
                __syncthreads();
                const int32_t jn = min(NT, q1 - tq);
#pragma unroll 2
                for (int32_t t = 0; t < jn; t++) {
                    const float4 q = tile_x[t];
                    const int32_t ej = tile_e[t], idj = tile_id[t];
#pragma unroll
                    for (int k = 0; k < R; k++) {
                        const float s = pair_s<true, ORTHO>(rx[k], ry[k], rz[k], q.x, q.y, q.z, B);
                        const bool cand = EDGE ? (s >= E.s_lo && s < E.s_hi) : in_range(s, A.g);
                        const bool ok = cand && rpos[k] < r1 && (!self || tq + t > rpos[k]);
                        if (ok) {
                            // the earlier atom is the reference's atomIndex: its element first
                            const bool rfirst = ri[k] < idj;
                            int64_t b;
                            if (EDGE) {
                                bool re64 = false;
                                b = rfirst ? edge_bin<true>(s, rx[k], ry[k], rz[k], q.x, q.y, q.z, B, E, re64)
                                           : edge_bin<true>(s, q.x, q.y, q.z, rx[k], ry[k], rz[k], B, E, re64);
                                nre += re64;
                                if (b < 0) continue;
                            } else {
                                b = bin_of(s, A.g);
                            }
                            const int32_t ea = rfirst ? re[k] : ej, eb = rfirst ? ej : re[k];
                            const int64_t flat = ((int64_t)ea * A.nel + eb) * A.g.H + b;
                            const bool intra = __float_as_int(q.w) == rm[k];
                            if (flat >= A.size) {
                                lost++;
                            } else if (LDS_HIST) {
                                atomicAdd(&lds_hist[(intra ? 0 : size) + (int32_t)flat], 1);
                            } else {
                                atomicAdd(&counts[(intra ? 0 : A.size) + flat], 1);
                            }
                        }
                    }
                }
            }
        }
    }
    if (lost) atomicAdd(&lds_overflow, lost);
    if (EDGE && nre) atomicAdd(&lds_rechecked, (unsigned long long)nre);
    __syncthreads();
    if (LDS_HIST) {
        int32_t *dst = partial + (int64_t)blockIdx.x * 2 * size;
        for (int32_t s = tid; s < 2 * size; s += NT) dst[s] = lds_hist[s];
    }
    if (tid == 0 && lds_overflow) atomicAdd(&counts[2 * A.size], lds_overflow);
    if (EDGE && tid == 0 && lds_rechecked)
        atomicAdd(reinterpret_cast<unsigned long long *>(&counts[2 * A.size + 2]), lds_rechecked);
}

// counts[s] += sum over the workgroup partials p in this split of partial[p][s]
__global__ void __launch_bounds__(256) k_sum_partials(const int32_t *__restrict__ partial, int32_t nparts,
                                                      int64_t len, int32_t per_split,
                                                      int32_t *__restrict__ counts) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= len) return;
    const int32_t p0 = blockIdx.y * per_split;
    const int32_t p1 = min(nparts, p0 + per_split);
    int32_t acc = 0;
    for (int32_t p = p0; p < p1; p++) acc += partial[(int64_t)p * len + s];
    if (acc) atomicAdd(&counts[s], acc);
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
    hipError_t e = pool_malloc((void **)&sc.p, sizeof(int32_t) * (2 * size + 1), st);
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
static int read_basis(const float *basis, int isPBC, hipStream_t st, Basis &B, bool &ortho) {
    ortho = false;
    for (int q = 0; q < 9; q++) B.b[q] = 0.f;
    if (!isPBC) return FRMC_OK;
    FRMC_CHECK_HIP(hipMemcpyAsync(B.b, basis, sizeof(B.b), hipMemcpyDeviceToHost, st));
    FRMC_CHECK_HIP(hipStreamSynchronize(st));
    const float *b = B.b;
    ortho = b[1] == 0.f && b[2] == 0.f && b[3] == 0.f && b[5] == 0.f && b[6] == 0.f && b[7] == 0.f;
    return FRMC_OK;
}

template <int NT, int R, bool EDGE>
static int launch_full_r(const float *coords, int32_t n, const Basis &B, int isPBC, bool ortho,
                         const int32_t *mol, const int32_t *elem, const HistArgs &A, FullSched S, const Edge64 &E,
                         int64_t blocks, bool lds, int32_t *partial, int32_t *counts, hipStream_t st) {
    const size_t shm = lds ? sizeof(int32_t) * 2 * A.size : 0;
#define FRMC_FULL_LAUNCH(P, O, L)                                                                               \
    do {                                                                                                        \
        if (shm > 48 * 1024)                                                                                    \
            (void)hipFuncSetAttribute((const void *)k_full_pairs<NT, R, P, O, L, EDGE>,                         \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);                    \
        k_full_pairs<NT, R, P, O, L, EDGE><<<(unsigned)blocks, NT, shm, st>>>(coords, n, B, mol, elem, A, S, E,\
                                                                              partial, counts);                 \
    } while (0)
    if (isPBC) {
        if (ortho) { if (lds) FRMC_FULL_LAUNCH(true, true, true); else FRMC_FULL_LAUNCH(true, true, false); }
        else       { if (lds) FRMC_FULL_LAUNCH(true, false, true); else FRMC_FULL_LAUNCH(true, false, false); }
    } else {
        if (lds) FRMC_FULL_LAUNCH(false, false, true); else FRMC_FULL_LAUNCH(false, false, false);
    }
#undef FRMC_FULL_LAUNCH
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

static FullSched full_sched(int32_t n, int32_t nt, int32_t R, int64_t blocks) {
    FullSched S;
    S.R = R;
    S.nct = (n + nt - 1) / nt;
    S.nrb = (n + R * nt - 1) / (R * nt);
    S.items = row_block_start(S.nrb, S.nct, R);
    if (blocks > S.items) blocks = S.items;
    S.per_block = (S.items + blocks - 1) / blocks;
    return S;
}

// ------------------------------------------------ full-histogram schedules

// A schedule of the full histogram: workgroup size (= column tile), row atoms
// per thread and resident workgroups per CU.  Every schedule gives the same
// counts (integer atomics), so the choice only changes the time.
struct FullChoice {
    int32_t nt = 256, R = 2, per_cu = 2;
    bool cells = false;  // the cell-list kernel (R = 1024 / nt rows per thread per pass)
};

// Cells per axis for pairs up to maxR apart in the periodic cell B: each
// axis gets floor(h / w) cells, h the spacing of the cell's lattice planes
// across it and w = maxR (1 + 1e-5) + 0.01 A (the margin covers the float32
// rounding of a pair's difference, image shift and norm).  False when an
// axis would have fewer than 3 cells (neighbour offsets -1, 0, +1 must name
// distinct cells) or the grid has fewer than 125 cells (too few to prune).
static bool cell_grid(const Basis &B, double maxR, int32_t m[3]) {
    double b[9];
    for (int q = 0; q < 9; q++) b[q] = B.b[q];
    const double vol = fabs(b[0] * (b[4] * b[8] - b[5] * b[7]) - b[1] * (b[3] * b[8] - b[5] * b[6]) +
                            b[2] * (b[3] * b[7] - b[4] * b[6]));
    const double w = maxR * (1.0 + 1e-5) + 0.01;
    for (int a = 0; a < 3; a++) {
        const double *u = b + 3 * ((a + 1) % 3), *v = b + 3 * ((a + 2) % 3);
        const double cx = u[1] * v[2] - u[2] * v[1], cy = u[2] * v[0] - u[0] * v[2], cz = u[0] * v[1] - u[1] * v[0];
        const double h = vol / sqrt(cx * cx + cy * cy + cz * cz);
        const double f = floor(h / w);
        if (!(f >= 3.0)) return false;
        m[a] = (int32_t)std::min(f, 64.0);
    }
    return (int64_t)m[0] * m[1] * m[2] >= 125;
}

constexpr int32_t kMaxPerCu = 16;  // workgroups per CU a schedule may ask for

// resident workgroups per CU for workgroup size nt: LDS (160 KB per CU, the
// private bins plus the staged tile) and 2,048 threads per CU, at most 8
static int32_t max_per_cu(size_t lds_bytes, bool lds, int32_t nt) {
    int32_t by_lds = lds ? (int32_t)((156 * 1024) / (lds_bytes + 20 * (size_t)nt + 2048)) : 8;
    int32_t p = by_lds < 2048 / nt ? by_lds : 2048 / nt;
    return p < 1 ? 1 : (p > 8 ? 8 : p);
}

// the fixed rule used when tuning is off: 256 threads, the residency limit,
// two row atoms per thread when that still leaves two items per workgroup
static FullChoice heuristic_choice(int32_t n, size_t lds_bytes, bool lds, int ncu) {
    FullChoice C;
    C.nt = 256;
    C.per_cu = max_per_cu(lds_bytes, lds, C.nt);
    C.R = full_sched(n, C.nt, 2, (int64_t)ncu * C.per_cu).items < 2 * (int64_t)ncu * C.per_cu ? 1 : 2;
    return C;
}

static int g_autotune = 1;                 // frmc_set_autotune
static FullChoice g_forced;                // frmc_force_full_schedule (R == 0: none)
static bool g_forced_on = false;
static thread_local FullChoice g_last;     // frmc_last_full_schedule (per thread)
static thread_local int64_t g_last_blocks = 0;
static thread_local int g_last_tuned = 0;
static thread_local int64_t g_last_evaluated = -1;  // pairs evaluated by the last call (-1: not measured)
static thread_local int64_t g_cells_evaluated = -1;  // set by run_cells when timing is on
static std::mutex g_tune_mu;
// shape class -> tuned schedule.  Class: device, floor(log2 n), bins
// (2*nel*nel*H), boundary kind (0 none, 1 triclinic, 2 orthorhombic), EDGE.
static std::map<std::tuple<int, int, int64_t, int, int>, FullChoice> g_tuned;

// One full-histogram pass with schedule C: counts += all pairs' bins.
static int run_cells(const FullChoice &C, const float *coords, int32_t n, const Basis &B, bool ortho,
                     const int32_t *mol, const int32_t *elem, const HistArgs &A, const Edge64 *E, int ncu,
                     int32_t *counts, hipStream_t st, int64_t *nblocks);

static double range_of(const HistArgs &A, const Edge64 *E) {
    return E ? sqrt((double)E->s_hi) : (double)A.g.maxD;
}

static int run_full(const FullChoice &C, const float *coords, int32_t n, const Basis &B, int isPBC, bool ortho,
                    const int32_t *mol, const int32_t *elem, const HistArgs &A, const Edge64 *E, int ncu,
                    int32_t *counts, hipStream_t st, int64_t *nblocks) {
    if (C.cells) return run_cells(C, coords, n, B, ortho, mol, elem, A, E, ncu, counts, st, nblocks);
    const size_t lds_bytes = sizeof(int32_t) * 2 * A.size;
    const bool lds = lds_bytes <= kMaxLdsHistBytes;
    const FullSched S = full_sched(n, C.nt, C.R, (int64_t)ncu * C.per_cu);
    const int64_t blocks = (S.items + S.per_block - 1) / S.per_block;
    if (nblocks) *nblocks = blocks;
    int32_t *partial = nullptr;
    if (lds) {
        hipError_t e = pool_malloc((void **)&partial, lds_bytes * blocks, st);
        if (e != hipSuccess) {
            set_error("hipMallocAsync of %lld bytes of partial histograms failed: %s",
                      (long long)(lds_bytes * blocks), hipGetErrorString(e));
            return FRMC_ENOMEM;
        }
    }
    Edge64 none;
    memset(&none, 0, sizeof(none));
    const Edge64 &EE = E ? *E : none;
    int rc = FRMC_EINVAL;
#define FRMC_FULL_CASE(NT_, R_)                                                                                \
    if (C.nt == NT_ && C.R == R_)                                                                              \
        rc = E ? launch_full_r<NT_, R_, true>(coords, n, B, isPBC, ortho, mol, elem, A, S, EE, blocks, lds,    \
                                              partial, counts, st)                                             \
               : launch_full_r<NT_, R_, false>(coords, n, B, isPBC, ortho, mol, elem, A, S, EE, blocks, lds,   \
                                               partial, counts, st);
    FRMC_FULL_CASE(256, 1)
    FRMC_FULL_CASE(256, 2)
    FRMC_FULL_CASE(256, 4)
    FRMC_FULL_CASE(512, 1)
    FRMC_FULL_CASE(512, 2)
    FRMC_FULL_CASE(512, 4)
    FRMC_FULL_CASE(1024, 1)
    FRMC_FULL_CASE(1024, 2)
    FRMC_FULL_CASE(1024, 4)
#undef FRMC_FULL_CASE
    if (rc == FRMC_EINVAL) set_error("no full-histogram kernel for %d threads and %d row atoms", C.nt, C.R);
    if (rc == FRMC_OK && lds) {
        const int32_t per_split = 16;
        const int64_t len = 2 * A.size;
        dim3 grid(blocks_for(len, 256), (unsigned)((blocks + per_split - 1) / per_split));
        k_sum_partials<<<grid, 256, 0, st>>>(partial, (int32_t)blocks, len, per_split, counts);
        hipError_t le = hipGetLastError();
        if (le != hipSuccess) {
            set_error("k_sum_partials launch failed: %s", hipGetErrorString(le));
            rc = FRMC_EHIP;
        }
    }
    if (partial) (void)hipFreeAsync(partial, st);
    return rc;
}

// The cell-list pass: sort the atoms into the cell grid (count, scan,
// scatter), then k_full_cells over (cell, offset) items and the partials'
// sum.  The caller checked cell_grid().
static int run_cells(const FullChoice &C, const float *coords, int32_t n, const Basis &B, bool ortho,
                     const int32_t *mol, const int32_t *elem, const HistArgs &A, const Edge64 *E, int ncu,
                     int32_t *counts, hipStream_t st, int64_t *nblocks) {
    int32_t m[3];
    if (!cell_grid(B, range_of(A, E), m)) {
        set_error("cell-list schedule on a cell too small for three cells per axis");
        return FRMC_EINVAL;
    }
    CellGrid G;
    G.m0 = m[0];
    G.m1 = m[1];
    G.m2 = m[2];
    G.ncell = (int64_t)m[0] * m[1] * m[2];
    G.items = 14 * G.ncell;
    int64_t blocks = std::min<int64_t>(G.items, (int64_t)ncu * C.per_cu);
    G.per_block = (G.items + blocks - 1) / blocks;
    blocks = (G.items + G.per_block - 1) / G.per_block;
    if (nblocks) *nblocks = blocks;
    const size_t lds_bytes = sizeof(int32_t) * 2 * A.size;
    const bool lds = lds_bytes <= kMaxLdsHistBytes;
    // scratch: cell of each atom, cell counts / fill cursors / starts, the
    // sorted atoms, then the workgroup partials
    const size_t nn = (size_t)(n > 0 ? n : 1), nc = (size_t)G.ncell;
    const size_t off_cnt = 16 * nn, off_fill = off_cnt + 4 * nc, off_start = off_fill + 4 * nc,
                 off_se = (off_start + 4 * (nc + 1) + 15) & ~(size_t)15, off_sid = off_se + 4 * nn,
                 off_cid = off_sid + 4 * nn, off_part = (off_cid + 4 * nn + 255) & ~(size_t)255;
    const size_t bytes = off_part + (lds ? lds_bytes * (size_t)blocks : 0);
    char *scr = nullptr;
    hipError_t e = pool_malloc((void **)&scr, bytes, st);
    if (e != hipSuccess) {
        set_error("hipMallocAsync of %zu bytes for the cell-list histogram failed: %s", bytes, hipGetErrorString(e));
        return FRMC_ENOMEM;
    }
    float4 *sx = (float4 *)scr;
    int32_t *cnt = (int32_t *)(scr + off_cnt), *fill = (int32_t *)(scr + off_fill),
            *cstart = (int32_t *)(scr + off_start), *se = (int32_t *)(scr + off_se), *sid = (int32_t *)(scr + off_sid),
            *cid = (int32_t *)(scr + off_cid), *partial = lds ? (int32_t *)(scr + off_part) : nullptr;
    k_fill_i32<<<blocks_for((int64_t)nc, 256), 256, 0, st>>>(cnt, (int64_t)nc, 0);
    k_cell_count<<<blocks_for(n, 256), 256, 0, st>>>(coords, n, G, cid, cnt);
    k_cell_scan<<<1, 1024, 0, st>>>(cnt, G.ncell, cstart, fill);
    k_cell_scatter<<<blocks_for(n, 256), 256, 0, st>>>(coords, n, mol, elem, cid, fill, sx, se, sid);
    hipError_t le = hipGetLastError();
    g_cells_evaluated = -1;
    if (le == hipSuccess && timing_enabled()) {
        // pairs the items visit: n_c (n_c - 1) / 2 within a cell, n_c n_c' for
        // each of its 13 visited neighbours (measurement only: synchronises)
        std::vector<int32_t> hs(nc + 1);
        le = hipMemcpyAsync(hs.data(), cstart, sizeof(int32_t) * (nc + 1), hipMemcpyDeviceToHost, st);
        if (le == hipSuccess) le = hipStreamSynchronize(st);
        if (le == hipSuccess) {
            int64_t ev = 0;
            for (int32_t x = 0; x < m[0]; x++)
                for (int32_t y = 0; y < m[1]; y++)
                    for (int32_t z = 0; z < m[2]; z++) {
                        const int64_t c = ((int64_t)x * m[1] + y) * m[2] + z;
                        const int64_t nc0 = hs[c + 1] - hs[c];
                        ev += nc0 * (nc0 - 1) / 2;
                        for (int32_t o = 1; o < 14; o++) {
                            const int32_t dx = o <= 9 ? 1 : 0, dy = o <= 9 ? (o - 1) / 3 - 1 : (o <= 12 ? 1 : 0),
                                          dz = o <= 9 ? (o - 1) % 3 - 1 : (o <= 12 ? o - 11 : 1);
                            const int64_t c2 = ((int64_t)((x + dx + m[0]) % m[0]) * m[1] + (y + dy + m[1]) % m[1]) *
                                                   m[2] + (z + dz + m[2]) % m[2];
                            ev += nc0 * (hs[c2 + 1] - hs[c2]);
                        }
                    }
            g_cells_evaluated = ev;
        }
    }
    Edge64 none;
    memset(&none, 0, sizeof(none));
    const Edge64 &EE = E ? *E : none;
    const size_t shm = lds ? lds_bytes : 0;
#define FRMC_CELLS_LAUNCH(NT_, O, L, ED)                                                                        \
    do {                                                                                                        \
        if (shm > 48 * 1024)                                                                                    \
            (void)hipFuncSetAttribute((const void *)k_full_cells<NT_, 1024 / NT_, O, L, ED>,                    \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);                    \
        k_full_cells<NT_, 1024 / NT_, O, L, ED><<<(unsigned)blocks, NT_, shm, st>>>(sx, se, sid, cstart, G, B, \
                                                                                    A, EE, partial, counts);    \
    } while (0)
#define FRMC_CELLS_NT(NT_)                                                                                      \
    do {                                                                                                        \
        if (ortho) {                                                                                            \
            if (lds) { if (E) FRMC_CELLS_LAUNCH(NT_, true, true, true); else FRMC_CELLS_LAUNCH(NT_, true, true, false); } \
            else     { if (E) FRMC_CELLS_LAUNCH(NT_, true, false, true); else FRMC_CELLS_LAUNCH(NT_, true, false, false); } \
        } else {                                                                                                \
            if (lds) { if (E) FRMC_CELLS_LAUNCH(NT_, false, true, true); else FRMC_CELLS_LAUNCH(NT_, false, true, false); } \
            else     { if (E) FRMC_CELLS_LAUNCH(NT_, false, false, true); else FRMC_CELLS_LAUNCH(NT_, false, false, false); } \
        }                                                                                                       \
    } while (0)
    int rc = FRMC_OK;
    if (le == hipSuccess) {
        if (C.nt == 256) FRMC_CELLS_NT(256);
        else if (C.nt == 512) FRMC_CELLS_NT(512);
        else rc = FRMC_EINVAL;
        le = hipGetLastError();
    }
#undef FRMC_CELLS_NT
#undef FRMC_CELLS_LAUNCH
    if (rc == FRMC_EINVAL) set_error("no cell-list kernel for %d threads", C.nt);
    if (rc == FRMC_OK && le == hipSuccess && lds) {
        const int32_t per_split = 16;
        const int64_t len = 2 * A.size;
        dim3 grid(blocks_for(len, 256), (unsigned)((blocks + per_split - 1) / per_split));
        k_sum_partials<<<grid, 256, 0, st>>>(partial, (int32_t)blocks, len, per_split, counts);
        le = hipGetLastError();
    }
    if (rc == FRMC_OK && le != hipSuccess) {
        set_error("cell-list histogram launch failed: %s", hipGetErrorString(le));
        rc = FRMC_EHIP;
    }
    (void)hipFreeAsync(scr, st);
    return rc;
}

// Times every candidate schedule on the caller's inputs (into scratch
// counts, so the caller's counts are untouched): each candidate runs twice
// between HIP events on the call stream and its faster run is kept.
static int tune_full(const float *coords, int32_t n, const Basis &B, int isPBC, bool ortho, const int32_t *mol,
                     const int32_t *elem, const HistArgs &A, const Edge64 *E, int ncu, hipStream_t st,
                     FullChoice &best) {
    const size_t lds_bytes = sizeof(int32_t) * 2 * A.size;
    const bool lds = lds_bytes <= kMaxLdsHistBytes;
    std::vector<FullChoice> cand;
    const FullChoice H0 = heuristic_choice(n, lds_bytes, lds, ncu);
    cand.push_back(H0);
    // workgroup sizes x row atoms x (1, 2, 4) times the resident workgroups
    // per CU: more workgroups than fit at once shorten the tail of the grid
    const int32_t nts[3] = {256, 512, 1024}, Rs[3] = {1, 2, 4};
    for (int32_t nt : nts) {
        const int32_t pmax = max_per_cu(lds_bytes, lds, nt);
        for (int32_t R : Rs)
            for (int32_t pc = pmax; pc <= 4 * pmax && pc <= kMaxPerCu; pc *= 2) {
                FullChoice C;
                C.nt = nt;
                C.R = R;
                C.per_cu = pc;
                // a schedule with fewer items than CUs leaves the chip idle
                if (full_sched(n, nt, R, 1).items < ncu) continue;
                bool dup = false;
                for (const FullChoice &D : cand) dup |= !D.cells && D.nt == nt && D.R == R && D.per_cu == pc;
                if (!dup) cand.push_back(C);
            }
    }
    // the cell-list kernel, when the cell is large enough against the range
    int32_t mg[3];
    if (isPBC && cell_grid(B, range_of(A, E), mg)) {
        const int32_t cnts[2] = {256, 512};
        for (int32_t nt : cnts) {
            const int32_t pmax = max_per_cu(lds_bytes + 4 * (size_t)nt, lds, nt);
            for (int32_t pc = pmax; pc <= 2 * pmax && pc <= kMaxPerCu; pc *= 2) {
                FullChoice C;
                C.nt = nt;
                C.R = 1024 / nt;
                C.per_cu = pc;
                C.cells = true;
                cand.push_back(C);
            }
        }
    }
    const int64_t len = 2 * A.size + 4;
    int32_t *scratch = nullptr;
    hipError_t e = pool_malloc((void **)&scratch, sizeof(int32_t) * len, st);
    if (e != hipSuccess) {
        set_error("hipMallocAsync of %lld tuning counts failed: %s", (long long)len, hipGetErrorString(e));
        return FRMC_ENOMEM;
    }
    const size_t nc = cand.size();
    std::vector<hipEvent_t> ev(4 * nc, nullptr);
    int rc = FRMC_OK;
    for (size_t q = 0; q < ev.size() && rc == FRMC_OK; q++)
        if (hipEventCreate(&ev[q]) != hipSuccess) {
            set_error("hipEventCreate failed");
            rc = FRMC_EHIP;
        }
    for (size_t c = 0; c < nc && rc == FRMC_OK; c++)
        for (int rep = 0; rep < 2 && rc == FRMC_OK; rep++) {
            k_fill_i32<<<64, 256, 0, st>>>(scratch, len, 0);
            if (hipEventRecord(ev[4 * c + 2 * rep], st) != hipSuccess) rc = FRMC_EHIP;
            if (rc == FRMC_OK)
                rc = run_full(cand[c], coords, n, B, isPBC, ortho, mol, elem, A, E, ncu, scratch, st, nullptr);
            if (rc == FRMC_OK && hipEventRecord(ev[4 * c + 2 * rep + 1], st) != hipSuccess) rc = FRMC_EHIP;
            if (rc == FRMC_EHIP) set_error("hipEventRecord failed while tuning");
        }
    if (rc == FRMC_OK && hipStreamSynchronize(st) != hipSuccess) {
        set_error("stream failed while tuning the full histogram");
        rc = FRMC_EHIP;
    }
    float best_ms = 0.f;
    best = H0;
    for (size_t c = 0; c < nc && rc == FRMC_OK; c++) {
        float m0 = 0.f, m1 = 0.f;
        if (hipEventElapsedTime(&m0, ev[4 * c], ev[4 * c + 1]) != hipSuccess ||
            hipEventElapsedTime(&m1, ev[4 * c + 2], ev[4 * c + 3]) != hipSuccess) {
            set_error("hipEventElapsedTime failed while tuning");
            rc = FRMC_EHIP;
            break;
        }
        const float m = m0 < m1 ? m0 : m1;
        if (c == 0 || m < best_ms) {
            best_ms = m;
            best = cand[c];
        }
    }
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    (void)hipFreeAsync(scratch, st);
    return rc;
}

// The full-histogram schedule: the forced one if set, else the tuned one for
// the shape class (tuned on the first call of the class), else the fixed rule.
static int choose_full(const float *coords, int32_t n, const Basis &B, int isPBC, bool ortho,
                       const int32_t *mol, const int32_t *elem, const HistArgs &A, const Edge64 *E, int dev,
                       int ncu, hipStream_t st, FullChoice &C, int &tuned) {
    const size_t lds_bytes = sizeof(int32_t) * 2 * A.size;
    const bool lds = lds_bytes <= kMaxLdsHistBytes;
    tuned = 0;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (g_forced_on) {
        C = g_forced;
        return FRMC_OK;
    }
    if (!g_autotune) {
        C = heuristic_choice(n, lds_bytes, lds, ncu);
        return FRMC_OK;
    }
    int lg = 0;
    while (lg < 30 && ((int64_t)1 << (lg + 1)) <= n) lg++;
    const auto key = std::make_tuple(dev, lg, A.size, isPBC ? (ortho ? 2 : 1) : 0, E ? 1 : 0);
    auto it = g_tuned.find(key);
    if (it == g_tuned.end()) {
        if (int rc = tune_full(coords, n, B, isPBC, ortho, mol, elem, A, E, ncu, st, C)) return rc;
        g_tuned[key] = C;
    } else {
        C = it->second;
    }
    tuned = 1;
    return FRMC_OK;
}

static int launch_full_timed(const float *coords, int32_t n, const Basis &B, int isPBC, bool ortho,
                             const int32_t *mol, const int32_t *elem, const HistArgs &A, const Edge64 *E,
                             int32_t *counts, hipStream_t st);

// All unordered pairs into counts ([2][nel][nel][H] + overflow, already
// zeroed or holding counts to add to).  Workgroups count into LDS-private
// bins and write them to per-workgroup partials, which k_sum_partials adds
// into counts (no contention on the global bins); histograms too large for
// LDS count straight into global memory with atomics.
int launch_full(const float *coords, int32_t n, const Basis &B, int isPBC, bool ortho, const int32_t *mol,
                const int32_t *elem, const HistArgs &A, int32_t *counts, hipStream_t st) {
    return launch_full_timed(coords, n, B, isPBC, ortho, mol, elem, A, nullptr, counts, st);
}

static int launch_full_timed(const float *coords, int32_t n, const Basis &B, int isPBC, bool ortho,
                             const int32_t *mol, const int32_t *elem, const HistArgs &A, const Edge64 *E,
                             int32_t *counts, hipStream_t st) {
    if (n < 2) return FRMC_OK;
    int dev = 0, ncu = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    FullChoice C;
    int tuned = 0;
    if (int rc = choose_full(coords, n, B, isPBC, ortho, mol, elem, A, E, dev, ncu, st, C, tuned)) return rc;
    hipEvent_t a = nullptr, b = nullptr;
    const bool timing = timing_enabled();
    if (timing) {
        FRMC_CHECK_HIP(hipEventCreate(&a));
        FRMC_CHECK_HIP(hipEventCreate(&b));
        FRMC_CHECK_HIP(hipEventRecord(a, st));
    }
    int64_t blocks = 0;
    int rc = run_full(C, coords, n, B, isPBC, ortho, mol, elem, A, E, ncu, counts, st, &blocks);
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
    g_last = C;
    g_last_blocks = blocks;
    g_last_tuned = tuned;
    g_last_evaluated = C.cells ? g_cells_evaluated : (int64_t)n * (n - 1) / 2;
    return rc;
}

// largest |x| of n floats, as the bit pattern of a non-negative float (NaN
// counts as infinity)
__global__ void __launch_bounds__(256) k_absmax(const float *__restrict__ x, int64_t n, uint32_t *out) {
    float m = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float a = fabsf(x[i]);
        m = (a != a) ? __int_as_float(0x7f800000) : fmaxf(m, a);
    }
    atomicMax(out, __float_as_uint(m));
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
    Basis hB;
    if ((rc = read_basis(basis, isPBC, st, hB, ortho))) return rc;
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
    Basis hB;
    if ((rc = read_basis(basis, isPBC, st, hB, ortho))) return rc;
    Scratch sc;
    if ((rc = alloc_counts(sc, A.size, st))) return rc;
    if ((rc = launch_full(coords, n, hB, isPBC, ortho, mol, elem, A, sc.p, st))) return rc;
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
        hipError_t e = pool_malloc((void **)&idx, sizeof(int32_t) * k, st);
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
    Basis hB;
    if ((rc = read_basis(basis, isPBC, st, hB, ortho))) return rc;
    k_fill_i32<<<1024, 256, 0, st>>>(counts, 2 * A.size + 1, 0);
    FRMC_CHECK_LAUNCH();
    return launch_full(coords, n, hB, isPBC, ortho, mol, elem, A, counts, st);
}

int frmc_full_pairs_counts_f64(const float *coords, int32_t n, const float *basis, int isPBC, const int32_t *mol,
                               const int32_t *elem, int32_t nel, float minD, float maxD, float bin, int32_t H,
                               int32_t *counts, int64_t *rechecked, void *stream) {
    HistArgs A;
    int rc = make_args(nel, H, minD, maxD, bin, A);
    if (rc) return rc;
    FRMC_REQUIRE(n >= 0 && counts, "bad arguments");
    FRMC_REQUIRE(n == 0 || (coords && mol && elem), "NULL pointer argument");
    FRMC_REQUIRE(!isPBC || basis, "basis is NULL");
    hipStream_t st = as_stream(stream);
    bool ortho = false;
    Basis hB;
    if ((rc = read_basis(basis, isPBC, st, hB, ortho))) return rc;
    // largest coordinate magnitude bounds the float32 difference error
    float amax = 0.0f;
    if (n > 0) {
        uint32_t *d_m = nullptr;
        hipError_t e = pool_malloc((void **)&d_m, sizeof(uint32_t), st);
        if (e != hipSuccess) {
            set_error("hipMallocAsync failed: %s", hipGetErrorString(e));
            return FRMC_ENOMEM;
        }
        k_fill_i32<<<1, 64, 0, st>>>((int32_t *)d_m, 1, 0);
        k_absmax<<<256, 256, 0, st>>>(coords, 3 * (int64_t)n, d_m);
        uint32_t bits = 0;
        e = hipMemcpyAsync(&bits, d_m, sizeof(bits), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        (void)hipFreeAsync(d_m, st);
        if (e != hipSuccess) {
            set_error("coordinate range reduction failed: %s", hipGetErrorString(e));
            return FRMC_EHIP;
        }
        memcpy(&amax, &bits, sizeof(amax));
        FRMC_REQUIRE(amax < 1e30f, "coordinates are not finite");
    }
    Edge64 E;
    E.mn = minD;
    E.mx = maxD;
    E.bn = bin;
    double rows = 0.0;
    for (int m = 0; m < 3; m++) {
        E.rowsum[m] = isPBC ? fabsf(hB.b[3 * m]) + fabsf(hB.b[3 * m + 1]) + fabsf(hB.b[3 * m + 2]) : 1.0f;
        rows += E.rowsum[m];
    }
    // widest per-pair error bound (twice the kernel's, for the window only)
    const double big = 2.0 * 8.0 * 5.9604644775390625e-08 * 1.001 * (rows * (2.0 * amax + 1.0) + maxD);
    const double lo = (double)minD - big, hi = (double)maxD + big;
    E.s_lo = lo > 0.0 ? nextafterf((float)(lo * lo), 0.0f) : 0.0f;
    E.s_hi = nextafterf((float)(hi * hi), INFINITY);
    k_fill_i32<<<1024, 256, 0, st>>>(counts, 2 * A.size + 4, 0);
    FRMC_CHECK_LAUNCH();
    rc = launch_full_timed(coords, n, hB, isPBC, ortho, mol, elem, A, &E, counts, st);
    if (rc) return rc;
    if (rechecked) {
        FRMC_CHECK_HIP(hipMemcpyAsync(rechecked, counts + 2 * A.size + 2, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        FRMC_CHECK_HIP(hipStreamSynchronize(st));
    }
    return FRMC_OK;
}

void frmc_set_autotune(int enable) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_autotune = enable != 0;
    if (enable < 0) g_tuned.clear();
}

int frmc_force_full_schedule(int32_t threads, int32_t rowAtoms, int32_t workgroupsPerCU) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (rowAtoms == 0) {
        g_forced_on = false;
        return FRMC_OK;
    }
    FRMC_REQUIRE(threads == 256 || threads == 512 || threads == 1024, "threads must be 256, 512 or 1024 (got %d)",
                 threads);
    FRMC_REQUIRE(rowAtoms == 1 || rowAtoms == 2 || rowAtoms == 4 || rowAtoms == -1,
                 "rowAtoms must be 1, 2, 4 or -1 (cell list) (got %d)", rowAtoms);
    FRMC_REQUIRE(rowAtoms != -1 || threads != 1024, "the cell-list kernel takes 256 or 512 threads");
    FRMC_REQUIRE(workgroupsPerCU >= 1 && workgroupsPerCU <= kMaxPerCu, "workgroupsPerCU must be in 1..%d (got %d)",
                 kMaxPerCu, workgroupsPerCU);
    g_forced.nt = threads;
    g_forced.cells = rowAtoms == -1;
    g_forced.R = g_forced.cells ? 1024 / threads : rowAtoms;
    g_forced.per_cu = workgroupsPerCU;
    g_forced_on = true;
    return FRMC_OK;
}

int frmc_last_full_evaluated(int64_t *pairs) {
    FRMC_REQUIRE(pairs, "NULL pointer argument");
    *pairs = g_last_evaluated;
    return FRMC_OK;
}

int frmc_last_full_schedule(int32_t *threads, int32_t *rowAtoms, int64_t *workgroups, int32_t *tuned) {
    FRMC_REQUIRE(threads && rowAtoms && workgroups && tuned, "NULL pointer argument");
    *threads = g_last.nt;
    *rowAtoms = g_last.cells ? -1 : g_last.R;
    *workgroups = g_last_blocks;
    *tuned = g_last_tuned;
    return FRMC_OK;
}

}  // extern "C"
