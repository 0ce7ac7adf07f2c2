// engine.hip -- device-resident RMC state and the speculative k-proposal step.
//
// State in HBM: box coordinates (n,3) float32, molecule/element index,
// group table (CSR), and per constraint the slot-exact int32 histograms
// counts[2][nel][nel][H] (intra, inter) plus their symmetrised pair-type sums
// base[T][H] (T = nel(nel+1)/2 pair types in sorted-element order), which is
// all G(r) consumes (PairDistributionConstraints.py:873-875).
//
// One window evaluates k proposals (group index + moved box coordinates)
// against the current state:
//   k_delta   fused "moved group vs all N" before/after binning for every
//             constraint grid at once: per proposal an int32 delta of the
//             pair-type histograms (LDS-private, merged once);
//   k_gr      G(r) of base+delta in the reference's float32 op order
//             (PairDistributionConstraints.py:854-895), PDC chi-square by
//             NumPy's pairwise summation;
//   k_sq      S(Q) = sequential-in-r float32 sum of G[r]*M[r,q], +1
//             (StructureFactorConstraints.py:772-773), 8 proposals per thread
//             so each M element read from L2 serves 8 sums;
//   k_finish  SFC chi-square (pairwise) and the total
//             np.sum([chi2_c]) (Engine.py:3026-3029).
// The accept rule (Engine.py:3308-3317) consumes a uniform draw only when
// chi-square got worse, so the host, which owns both MT19937 streams, walks
// the k results in order and commits the first accepted proposal:
//   k_commit_slots  slot-exact update data - before + after (M - F slot
//                   convention of PairDistributionConstraints.py:1053-1078),
//   k_commit_small  pair-type sums and coordinates.
#include <math.h>
#include <stdlib.h>

#include <algorithm>
#include <vector>

#include "frmc_device.h"
#include "frmc_host.h"

namespace frmc {

constexpr int kMaxC = 4;        // constraints per engine
constexpr int kMaxG = 64;       // atoms per group in the fused kernels
constexpr int kPB = 8;          // proposals per thread in k_sq
constexpr int kSqRows = 128;    // G rows staged per LDS pass in k_sq

struct DevC {
    int32_t kind;   // FRMC_PDC or FRMC_SFC
    Grid g;
    int32_t H, M;   // histogram bins, model points (H for PDC, nQ for SFC)
    int32_t hasW, hasSF;
    float sf;
    const float *w, *D, *shellV, *pref, *expv, *dataW, *sqm;
    const int32_t *leaves;  // (start,len) pairs of the pairwise-sum leaves of length M
    const int32_t *prog;    // postfix combine program over the leaves
    int32_t nleaves, nprog;
    int32_t *base;          // [T][H]
    int32_t *slot;          // [2][nel][nel][H] + overflow counter
    int32_t dnOff;          // offset in a proposal's delta block
};

struct View {
    int32_t n, nel, T, C, isPBC, ortho;
    Basis B;
    float *xyz;                    // (n,3) current box coordinates
    const int32_t *mol, *elem;
    const int32_t *goff, *gatoms;  // groups
    DevC c[kMaxC];
    int32_t dnStride;              // ints per proposal delta block
    int32_t Hmax, Mmax;
    // window buffers
    const int32_t *pgroup;  // [k]
    const int32_t *poff;    // [k+1] atom offsets into pxyz
    const float *pxyz;      // [sum g][3]
    int32_t *dn;            // [k][dnStride]
    float *G;               // [k][C][Hmax]
    float *sqv;             // [k][C][Mmax]
    float *chi2;            // [k][C]
    float *tse;             // [k]
};

__device__ __forceinline__ int32_t pair_type(int32_t a, int32_t b, int32_t nel) {
    int32_t lo = a < b ? a : b, hi = a < b ? b : a;
    return lo * (2 * nel - lo + 1) / 2 + (hi - lo);
}

template <bool PBC, bool ORTHO>
__device__ __forceinline__ float eng_pair_s(float ax, float ay, float az, float cx, float cy, float cz,
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

// Pair-type histogram index of an in-range pair counted in slot [ea][eb];
// reproduces the flat-index aliasing of binIndex == H.  Returns -1 when the
// flat index leaves the (nel,nel,H) array (dropped).
__device__ __forceinline__ int32_t type_bin_index(float s, const DevC &c, int32_t ea, int32_t eb, int32_t nel) {
    int32_t b = bin_of(s, c.g);
    if (b >= 0 && b < c.H) return pair_type(ea, eb, nel) * c.H + b;
    int64_t flat = ((int64_t)ea * nel + eb) * c.H + b;
    if (flat < 0 || flat >= (int64_t)nel * nel * c.H) return -1;
    int32_t a2 = (int32_t)(flat / ((int64_t)nel * c.H));
    int32_t b2 = (int32_t)((flat / c.H) % nel);
    return pair_type(a2, b2, nel) * c.H + (int32_t)(flat % c.H);
}

// ----------------------------------------------------------------- k_delta

template <bool PBC, bool ORTHO>
__global__ void __launch_bounds__(256) k_delta(View V, int32_t chunkLen, int32_t merge_atomic) {
    extern __shared__ int32_t lds[];  // V.dnStride ints
    __shared__ float g_old[kMaxG][3], g_new[kMaxG][3];
    __shared__ int32_t g_id[kMaxG], g_el[kMaxG];
    const int32_t p = blockIdx.y;
    const int tid = threadIdx.x;
    for (int32_t s = tid; s < V.dnStride; s += blockDim.x) lds[s] = 0;
    const int32_t gi = V.pgroup[p];
    const int32_t a0 = V.poff[p];
    const int32_t g = V.poff[p + 1] - a0;
    const int32_t *atoms = V.gatoms + V.goff[gi];
    for (int32_t a = tid; a < g; a += blockDim.x) {
        int32_t id = atoms[a];
        g_id[a] = id;
        g_el[a] = V.elem[id];
        g_old[a][0] = V.xyz[3 * (int64_t)id];
        g_old[a][1] = V.xyz[3 * (int64_t)id + 1];
        g_old[a][2] = V.xyz[3 * (int64_t)id + 2];
        g_new[a][0] = V.pxyz[3 * (int64_t)(a0 + a)];
        g_new[a][1] = V.pxyz[3 * (int64_t)(a0 + a) + 1];
        g_new[a][2] = V.pxyz[3 * (int64_t)(a0 + a) + 2];
    }
    __syncthreads();

    const int32_t j0 = blockIdx.x * chunkLen;
    const int32_t j1 = min(V.n, j0 + chunkLen);
    for (int32_t j = j0 + tid; j < j1; j += blockDim.x) {
        bool member = false;
        for (int32_t a = 0; a < g; a++) member |= (g_id[a] == j);
        if (member) continue;
        const float cx = V.xyz[3 * (int64_t)j], cy = V.xyz[3 * (int64_t)j + 1], cz = V.xyz[3 * (int64_t)j + 2];
        const int32_t ej = V.elem[j];
        for (int32_t a = 0; a < g; a++) {
            // before: M counts (atom a, other j) with diff = c[a]-c[j]; after: moved a
            float so = eng_pair_s<PBC, ORTHO>(g_old[a][0], g_old[a][1], g_old[a][2], cx, cy, cz, V.B);
            float sn = eng_pair_s<PBC, ORTHO>(g_new[a][0], g_new[a][1], g_new[a][2], cx, cy, cz, V.B);
            for (int32_t ci = 0; ci < V.C; ci++) {
                const DevC &c = V.c[ci];
                if (in_range(so, c.g)) {
                    int32_t ix = type_bin_index(so, c, g_el[a], ej, V.nel);
                    if (ix >= 0) atomicSub(&lds[c.dnOff + ix], 1);
                }
                if (in_range(sn, c.g)) {
                    int32_t ix = type_bin_index(sn, c, g_el[a], ej, V.nel);
                    if (ix >= 0) atomicAdd(&lds[c.dnOff + ix], 1);
                }
            }
        }
    }
    // group-internal pairs: M - F leaves each unordered pair once, in slot
    // [e(later position), e(earlier position)]
    if (blockIdx.x == 0 && g > 1) {
        const int32_t npairs = g * (g - 1) / 2;
        for (int32_t q = tid; q < npairs; q += blockDim.x) {
            // decode q -> (a > b)
            int32_t a = 1, rem = q;
            while (rem >= a) { rem -= a; a++; }
            const int32_t b = rem;
            float so = eng_pair_s<PBC, ORTHO>(g_old[a][0], g_old[a][1], g_old[a][2], g_old[b][0], g_old[b][1],
                                              g_old[b][2], V.B);
            float sn = eng_pair_s<PBC, ORTHO>(g_new[a][0], g_new[a][1], g_new[a][2], g_new[b][0], g_new[b][1],
                                              g_new[b][2], V.B);
            for (int32_t ci = 0; ci < V.C; ci++) {
                const DevC &c = V.c[ci];
                if (in_range(so, c.g)) {
                    int32_t ix = type_bin_index(so, c, g_el[a], g_el[b], V.nel);
                    if (ix >= 0) atomicSub(&lds[c.dnOff + ix], 1);
                }
                if (in_range(sn, c.g)) {
                    int32_t ix = type_bin_index(sn, c, g_el[a], g_el[b], V.nel);
                    if (ix >= 0) atomicAdd(&lds[c.dnOff + ix], 1);
                }
            }
        }
    }
    __syncthreads();
    int32_t *out = V.dn + (int64_t)p * V.dnStride;
    if (merge_atomic) {
        for (int32_t s = tid; s < V.dnStride; s += blockDim.x) {
            int32_t v = lds[s];
            if (v) atomicAdd(&out[s], v);
        }
    } else {
        for (int32_t s = tid; s < V.dnStride; s += blockDim.x) out[s] = lds[s];
    }
}

// ------------------------------------------------------ pairwise summation

// NumPy's pairwise float32 sum (add.reduce over a contiguous 1-D array):
// leaves of <= 128 elements use 8 strided accumulators, combined
// ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)), then the tail; leaves of < 8 are
// summed sequentially from 0.  The leaves are split as NumPy splits
// (n2 = n/2 rounded down to a multiple of 8) and recombined in the same tree
// by a postfix program computed on the host.
__device__ float leaf_sum(const float *a, int32_t n) {
    if (n < 8) {
        float r = 0.0f;
        for (int32_t i = 0; i < n; i++) r = r + a[i];
        return r;
    }
    float r[8];
#pragma unroll
    for (int j = 0; j < 8; j++) r[j] = a[j];
    int32_t i = 8;
    for (; i < n - (n % 8); i += 8) {
#pragma unroll
        for (int j = 0; j < 8; j++) r[j] = r[j] + a[i + j];
    }
    float res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; i++) res = res + a[i];
    return res;
}

// Block-cooperative: every thread calls it; returns the sum in thread 0.
// scratch needs nleaves floats of LDS.
__device__ float block_pairwise(const float *a, const DevC &c, float *scratch) {
    for (int32_t l = threadIdx.x; l < c.nleaves; l += blockDim.x)
        scratch[l] = leaf_sum(a + c.leaves[2 * l], c.leaves[2 * l + 1]);
    __syncthreads();
    float res = 0.0f;
    if (threadIdx.x == 0) {
        float st[40];
        int sp = 0;
        for (int32_t k = 0; k < c.nprog; k++) {
            int32_t op = c.prog[k];
            if (op >= 0) {
                st[sp++] = scratch[op];
            } else {
                float b = st[--sp];
                float x = st[--sp];
                st[sp++] = x + b;
            }
        }
        res = st[0];
    }
    __syncthreads();
    return res;
}

// -------------------------------------------------------------------- k_gr

// grid (k, C); G(r) of base + delta[p] for constraint c.
__global__ void __launch_bounds__(256) k_gr(View V) {
    extern __shared__ float lsq[];  // Hmax floats + leaf scratch
    const int32_t p = blockIdx.x, ci = blockIdx.y;
    const DevC &c = V.c[ci];
    const int32_t *dn = V.dn + (int64_t)p * V.dnStride + c.dnOff;
    float *Gout = V.G + ((int64_t)p * V.C + ci) * V.Hmax;
    for (int32_t r = threadIdx.x; r < c.H; r += blockDim.x) {
        float Gr = 0.0f;
        for (int32_t t = 0; t < V.T; t++) {
            // Gr += wij*nij/Dij  (nij integer-valued, exact in float32)
            float nij = (float)(c.base[t * c.H + r] + dn[t * c.H + r]);
            float term = c.w[t] * nij;
            term = term / c.D[t];
            Gr = Gr + term;
        }
        Gr = Gr / c.shellV[r];                // Gr /= shellVolumes
        float G = c.pref[r] * (Gr - 1.0f);    // (4*PI*r*rho0)*(Gr-1)
        if (c.kind == FRMC_PDC) {
            if (c.hasSF) G = G * c.sf;
            float diff = c.expv[r] - G;
            float v = diff * diff;
            if (c.hasW) v = c.dataW[r] * v;
            lsq[r] = v;
        } else {
            Gout[r] = G;
        }
    }
    if (c.kind == FRMC_PDC) {
        __syncthreads();
        float chi = block_pairwise(lsq, c, lsq + V.Hmax);
        if (threadIdx.x == 0) V.chi2[(int64_t)p * V.C + ci] = chi;
    }
}

// -------------------------------------------------------------------- k_sq

// grid (ceil(M/256), ceil(k/kPB), nSFC); thread = one q for kPB proposals.
__global__ void __launch_bounds__(256) k_sq(View V, int32_t k, const int32_t *__restrict__ sfc_ids) {
    __shared__ float Gs[kPB][kSqRows];
    const int32_t ci = sfc_ids[blockIdx.z];
    const DevC &c = V.c[ci];
    const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t p0 = blockIdx.y * kPB;
    float acc[kPB];
#pragma unroll
    for (int pb = 0; pb < kPB; pb++) acc[pb] = 0.0f;
    for (int32_t r0 = 0; r0 < c.H; r0 += kSqRows) {
        const int32_t rows = min(kSqRows, c.H - r0);
        __syncthreads();
        for (int32_t s = threadIdx.x; s < kPB * kSqRows; s += blockDim.x) {
            int32_t pb = s / kSqRows, rr = s % kSqRows;
            int32_t p = p0 + pb;
            Gs[pb][rr] = (p < k && rr < rows) ? V.G[((int64_t)p * V.C + ci) * V.Hmax + r0 + rr] : 0.0f;
        }
        __syncthreads();
        if (q < c.M) {
            for (int32_t rr = 0; rr < rows; rr++) {
                const float m = c.sqm[(int64_t)(r0 + rr) * c.M + q];
                if (r0 + rr == 0) {
#pragma unroll
                    for (int pb = 0; pb < kPB; pb++) acc[pb] = Gs[pb][rr] * m;
                } else {
#pragma unroll
                    for (int pb = 0; pb < kPB; pb++) {
                        float prod = Gs[pb][rr] * m;
                        acc[pb] = acc[pb] + prod;
                    }
                }
            }
        }
    }
    if (q >= c.M) return;
#pragma unroll
    for (int pb = 0; pb < kPB; pb++) {
        const int32_t p = p0 + pb;
        if (p >= k) break;
        float S = acc[pb] + 1.0f;
        if (c.hasSF) {
            float t = S - 1.0f;
            t = c.sf * t;
            S = t + 1.0f;
        }
        float diff = c.expv[q] - S;
        float v = diff * diff;
        if (c.hasW) v = c.dataW[q] * v;
        V.sqv[((int64_t)p * V.C + ci) * V.Mmax + q] = v;
    }
}

// ---------------------------------------------------------------- k_finish

// grid k: SFC chi-squares by pairwise summation, then the total.
__global__ void __launch_bounds__(256) k_finish(View V) {
    extern __shared__ float scratch[];
    const int32_t p = blockIdx.x;
    for (int32_t ci = 0; ci < V.C; ci++) {
        const DevC &c = V.c[ci];
        if (c.kind != FRMC_SFC) continue;
        float chi = block_pairwise(V.sqv + ((int64_t)p * V.C + ci) * V.Mmax, c, scratch);
        if (threadIdx.x == 0) V.chi2[(int64_t)p * V.C + ci] = chi;
    }
    if (threadIdx.x == 0) {
        // np.sum([SD_c / varianceSquared]) with varianceSquared == 1: fewer than
        // 8 terms -> sequential from 0
        float tse = 0.0f;
        for (int32_t ci = 0; ci < V.C; ci++) tse = tse + V.chi2[(int64_t)p * V.C + ci];
        V.tse[p] = tse;
    }
}

// ---------------------------------------------------------------- commits

template <bool PBC, bool ORTHO>
__global__ void __launch_bounds__(256) k_commit_slots(View V, int32_t p) {
    __shared__ float g_old[kMaxG][3], g_new[kMaxG][3];
    __shared__ int32_t g_id[kMaxG], g_el[kMaxG], g_mol[kMaxG];
    const int tid = threadIdx.x;
    const int32_t gi = V.pgroup[p];
    const int32_t a0 = V.poff[p];
    const int32_t g = V.poff[p + 1] - a0;
    const int32_t *atoms = V.gatoms + V.goff[gi];
    for (int32_t a = tid; a < g; a += blockDim.x) {
        int32_t id = atoms[a];
        g_id[a] = id;
        g_el[a] = V.elem[id];
        g_mol[a] = V.mol[id];
        for (int d = 0; d < 3; d++) {
            g_old[a][d] = V.xyz[3 * (int64_t)id + d];
            g_new[a][d] = V.pxyz[3 * (int64_t)(a0 + a) + d];
        }
    }
    __syncthreads();
    const int32_t j = blockIdx.x * blockDim.x + tid;
    if (j < V.n) {
        bool member = false;
        for (int32_t a = 0; a < g; a++) member |= (g_id[a] == j);
        if (!member) {
            const float cx = V.xyz[3 * (int64_t)j], cy = V.xyz[3 * (int64_t)j + 1], cz = V.xyz[3 * (int64_t)j + 2];
            const int32_t ej = V.elem[j], mj = V.mol[j];
            for (int32_t a = 0; a < g; a++) {
                float so = eng_pair_s<PBC, ORTHO>(g_old[a][0], g_old[a][1], g_old[a][2], cx, cy, cz, V.B);
                float sn = eng_pair_s<PBC, ORTHO>(g_new[a][0], g_new[a][1], g_new[a][2], cx, cy, cz, V.B);
                const bool intra = g_mol[a] == mj;
                for (int32_t ci = 0; ci < V.C; ci++) {
                    const DevC &c = V.c[ci];
                    const int64_t size = (int64_t)V.nel * V.nel * c.H;
                    int32_t *h = c.slot + (intra ? 0 : size);
                    if (in_range(so, c.g)) {
                        int64_t f = ((int64_t)g_el[a] * V.nel + ej) * c.H + bin_of(so, c.g);
                        if (f >= 0 && f < size) atomicSub(&h[f], 1); else atomicAdd(&c.slot[2 * size], 1);
                    }
                    if (in_range(sn, c.g)) {
                        int64_t f = ((int64_t)g_el[a] * V.nel + ej) * c.H + bin_of(sn, c.g);
                        if (f >= 0 && f < size) atomicAdd(&h[f], 1); else atomicAdd(&c.slot[2 * size], 1);
                    }
                }
            }
        }
    }
    if (blockIdx.x == 0 && g > 1) {
        const int32_t npairs = g * (g - 1) / 2;
        for (int32_t q = tid; q < npairs; q += blockDim.x) {
            int32_t a = 1, rem = q;
            while (rem >= a) { rem -= a; a++; }
            const int32_t b = rem;
            float so = eng_pair_s<PBC, ORTHO>(g_old[a][0], g_old[a][1], g_old[a][2], g_old[b][0], g_old[b][1],
                                              g_old[b][2], V.B);
            float sn = eng_pair_s<PBC, ORTHO>(g_new[a][0], g_new[a][1], g_new[a][2], g_new[b][0], g_new[b][1],
                                              g_new[b][2], V.B);
            const bool intra = g_mol[a] == g_mol[b];
            for (int32_t ci = 0; ci < V.C; ci++) {
                const DevC &c = V.c[ci];
                const int64_t size = (int64_t)V.nel * V.nel * c.H;
                int32_t *h = c.slot + (intra ? 0 : size);
                if (in_range(so, c.g)) {
                    int64_t f = ((int64_t)g_el[a] * V.nel + g_el[b]) * c.H + bin_of(so, c.g);
                    if (f >= 0 && f < size) atomicSub(&h[f], 1); else atomicAdd(&c.slot[2 * size], 1);
                }
                if (in_range(sn, c.g)) {
                    int64_t f = ((int64_t)g_el[a] * V.nel + g_el[b]) * c.H + bin_of(sn, c.g);
                    if (f >= 0 && f < size) atomicAdd(&h[f], 1); else atomicAdd(&c.slot[2 * size], 1);
                }
            }
        }
    }
}

// one block: pair-type sums += delta[p]; coordinates <- proposal p
__global__ void __launch_bounds__(256) k_commit_small(View V, int32_t p) {
    const int32_t *dn = V.dn + (int64_t)p * V.dnStride;
    for (int32_t ci = 0; ci < V.C; ci++) {
        const DevC &c = V.c[ci];
        for (int32_t s = threadIdx.x; s < V.T * c.H; s += blockDim.x) c.base[s] += dn[c.dnOff + s];
    }
    const int32_t gi = V.pgroup[p];
    const int32_t a0 = V.poff[p];
    const int32_t g = V.poff[p + 1] - a0;
    const int32_t *atoms = V.gatoms + V.goff[gi];
    for (int32_t a = threadIdx.x; a < g; a += blockDim.x) {
        int32_t id = atoms[a];
        for (int d = 0; d < 3; d++) V.xyz[3 * (int64_t)id + d] = V.pxyz[3 * (int64_t)(a0 + a) + d];
    }
}

// base[t][r] = symmetrised slot sums (intra+inter, both orders)
__global__ void k_symmetrise(int32_t *__restrict__ base, const int32_t *__restrict__ slot, int32_t nel, int32_t H) {
    const int32_t T = nel * (nel + 1) / 2;
    const int64_t size = (int64_t)nel * nel * H;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < (int64_t)T * H;
         s += (int64_t)gridDim.x * blockDim.x) {
        int32_t t = (int32_t)(s / H), r = (int32_t)(s % H);
        // decode t -> (i <= j)
        int32_t i = 0, rem = t;
        while (rem >= nel - i) { rem -= nel - i; i++; }
        int32_t j = i + rem;
        int64_t ij = ((int64_t)i * nel + j) * H + r, ji = ((int64_t)j * nel + i) * H + r;
        int32_t v;
        if (i == j) v = slot[ij] + slot[size + ij];
        else v = ((slot[ij] + slot[ji]) + slot[size + ij]) + slot[size + ji];
        base[s] = v;
    }
}

__global__ void k_fill_i32_e(int32_t *__restrict__ p, int64_t n, int32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ------------------------------------------------------------------- host

// NumPy pairwise-sum leaves and combine program for length n.
static void pairwise_plan(int32_t n, std::vector<int32_t> &leaves, std::vector<int32_t> &prog) {
    struct Rec {
        static void go(int32_t start, int32_t len, std::vector<int32_t> &lv, std::vector<int32_t> &pg) {
            if (len <= 128) {
                pg.push_back((int32_t)(lv.size() / 2));
                lv.push_back(start);
                lv.push_back(len);
                return;
            }
            int32_t n2 = len / 2;
            n2 -= n2 % 8;
            go(start, n2, lv, pg);
            go(start + n2, len - n2, lv, pg);
            pg.push_back(-1);
        }
    };
    leaves.clear();
    prog.clear();
    Rec::go(0, n, leaves, prog);
}

struct HostC {
    frmc_constraint_desc d;
    DevC dev;
    void *allocs[16];
    int nallocs = 0;
};

}  // namespace frmc

struct frmc_engine {
    hipStream_t st = nullptr;
    bool own_stream = false;
    int32_t n = 0, nel = 0, T = 0, C = 0, G = 0, isPBC = 0, ortho = 0, kmax = 0, gmax = 0, total_gatoms = 0;
    frmc::Basis B;
    float *xyz = nullptr, *dbasis = nullptr;
    int32_t *mol = nullptr, *elem = nullptr, *goff = nullptr, *gatoms = nullptr;
    std::vector<int32_t> h_goff, h_gatoms;
    frmc::HostC c[frmc::kMaxC];
    int32_t nsfc = 0;
    int32_t *sfc_ids = nullptr;
    int32_t dnStride = 0, Hmax = 0, Mmax = 0;
    // window buffers
    int32_t *pgroup = nullptr, *poff = nullptr;
    float *pxyz = nullptr;
    int32_t *dn = nullptr;
    float *Gb = nullptr, *sqv = nullptr, *chi2 = nullptr, *tse = nullptr;
    // pinned staging
    int32_t *h_pgroup = nullptr, *h_poff = nullptr;
    float *h_pxyz = nullptr, *h_out = nullptr;  // h_out: [k] tse then [k][C] chi2
    int32_t last_k = 0;
    // host mirror of the accepted state
    float cur_tse = 0.f;
    float cur_chi2[frmc::kMaxC] = {0};
    std::vector<float> h_real, h_box;  // (n,3) host mirrors used by the native generator
    float recip[9] = {0};
    bool have_real = false;
    std::vector<void *> allocs;
};

namespace frmc {

static View make_view(frmc_engine *e) {
    View V;
    V.n = e->n;
    V.nel = e->nel;
    V.T = e->T;
    V.C = e->C;
    V.isPBC = e->isPBC;
    V.ortho = e->ortho;
    V.B = e->B;
    V.xyz = e->xyz;
    V.mol = e->mol;
    V.elem = e->elem;
    V.goff = e->goff;
    V.gatoms = e->gatoms;
    for (int i = 0; i < e->C; i++) V.c[i] = e->c[i].dev;
    V.dnStride = e->dnStride;
    V.Hmax = e->Hmax;
    V.Mmax = e->Mmax;
    V.pgroup = e->pgroup;
    V.poff = e->poff;
    V.pxyz = e->pxyz;
    V.dn = e->dn;
    V.G = e->Gb;
    V.sqv = e->sqv;
    V.chi2 = e->chi2;
    V.tse = e->tse;
    return V;
}

template <typename T>
static hipError_t dalloc(frmc_engine *e, T **p, size_t count) {
    hipError_t err = hipMalloc((void **)p, sizeof(T) * (count ? count : 1));
    if (err == hipSuccess) e->allocs.push_back((void *)*p);
    return err;
}

template <typename T>
static hipError_t dupload(frmc_engine *e, T **p, const T *host, size_t count) {
    hipError_t err = dalloc(e, p, count);
    if (err != hipSuccess) return err;
    if (count) err = hipMemcpy((void *)*p, host, sizeof(T) * count, hipMemcpyHostToDevice);
    return err;
}

static void free_engine(frmc_engine *e) {
    if (!e) return;
    if (e->st) (void)hipStreamSynchronize(e->st);
    for (void *p : e->allocs) (void)hipFree(p);
    if (e->h_pgroup) (void)hipHostFree(e->h_pgroup);
    if (e->h_poff) (void)hipHostFree(e->h_poff);
    if (e->h_pxyz) (void)hipHostFree(e->h_pxyz);
    if (e->h_out) (void)hipHostFree(e->h_out);
    if (e->own_stream && e->st) (void)hipStreamDestroy(e->st);
    delete e;
}

// Launch the evaluation of the k proposals already in the window buffers.
static int launch_eval(frmc_engine *e, int32_t k) {
    View V = make_view(e);
    hipStream_t st = e->st;
    // delta: one chunk per proposal when n is small, otherwise split n so
    // that the grid holds at least ~2 blocks per CU
    int32_t nchunk = 1;
    if ((int64_t)k * 2 < 512) nchunk = (int32_t)std::min<int64_t>(std::max<int64_t>(1, 512 / (k > 0 ? k : 1)),
                                                                  std::max<int64_t>(1, e->n / 2048));
    int32_t chunkLen = (e->n + nchunk - 1) / nchunk;
    if (chunkLen < 1) chunkLen = 1;
    nchunk = (e->n + chunkLen - 1) / chunkLen;
    if (nchunk < 1) nchunk = 1;
    const int merge_atomic = nchunk > 1;
    if (merge_atomic) {
        k_fill_i32_e<<<blocks_for((int64_t)k * e->dnStride, 256) > 2048 ? 2048 : blocks_for((int64_t)k * e->dnStride, 256),
                       256, 0, st>>>(e->dn, (int64_t)k * e->dnStride, 0);
    }
    size_t shm = sizeof(int32_t) * e->dnStride;
    dim3 grid((unsigned)nchunk, (unsigned)k);
#define FRMC_DELTA(P, O)                                                                                     \
    do {                                                                                                     \
        if (shm > 48 * 1024)                                                                                 \
            (void)hipFuncSetAttribute((const void *)k_delta<P, O>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                      (int)shm);                                                             \
        k_delta<P, O><<<grid, 256, shm, st>>>(V, chunkLen, merge_atomic);                                     \
    } while (0)
    if (e->isPBC) {
        if (e->ortho) FRMC_DELTA(true, true); else FRMC_DELTA(true, false);
    } else {
        FRMC_DELTA(false, false);
    }
#undef FRMC_DELTA
    FRMC_CHECK_LAUNCH();
    int32_t maxLeaves = 1;
    for (int i = 0; i < e->C; i++) maxLeaves = std::max(maxLeaves, e->c[i].dev.nleaves);
    size_t gshm = sizeof(float) * (e->Hmax + maxLeaves);
    if (gshm > 48 * 1024) (void)hipFuncSetAttribute((const void *)k_gr, hipFuncAttributeMaxDynamicSharedMemorySize, (int)gshm);
    k_gr<<<dim3((unsigned)k, (unsigned)e->C), 256, gshm, st>>>(V);
    FRMC_CHECK_LAUNCH();
    if (e->nsfc > 0) {
        dim3 sg(blocks_for(e->Mmax, 256), (unsigned)((k + kPB - 1) / kPB), (unsigned)e->nsfc);
        k_sq<<<sg, 256, 0, st>>>(V, k, e->sfc_ids);
        FRMC_CHECK_LAUNCH();
    }
    k_finish<<<(unsigned)k, 256, sizeof(float) * maxLeaves, st>>>(V);
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

static int launch_commit(frmc_engine *e, int32_t p) {
    View V = make_view(e);
    hipStream_t st = e->st;
    unsigned g = blocks_for(e->n, 256);
    if (e->isPBC) {
        if (e->ortho) k_commit_slots<true, true><<<g, 256, 0, st>>>(V, p);
        else k_commit_slots<true, false><<<g, 256, 0, st>>>(V, p);
    } else {
        k_commit_slots<false, false><<<g, 256, 0, st>>>(V, p);
    }
    FRMC_CHECK_LAUNCH();
    k_commit_small<<<1, 256, 0, st>>>(V, p);
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

int launch_full(const float *coords, int32_t n, const float *basis, int isPBC, bool ortho, const int32_t *mol,
                const int32_t *elem, const HistArgs &A, int32_t *counts, hipStream_t st);

}  // namespace frmc

using namespace frmc;

// ----------------------------------------------------------------- C ABI

extern "C" {

int frmc_engine_create(frmc_engine **out, const frmc_engine_desc *d) {
    FRMC_REQUIRE(out && d, "NULL argument");
    *out = nullptr;
    FRMC_REQUIRE(d->n >= 2, "need at least 2 atoms (n=%d)", d->n);
    FRMC_REQUIRE(d->nel >= 1 && d->nel <= 64, "numberOfElements %d out of range [1,64]", d->nel);
    FRMC_REQUIRE(d->nconstraints >= 1 && d->nconstraints <= kMaxC, "1..%d constraints supported", kMaxC);
    FRMC_REQUIRE(d->ngroups >= 1, "no groups");
    FRMC_REQUIRE(d->kmax >= 1 && d->kmax <= 4096, "kmax %d out of range [1,4096]", d->kmax);
    FRMC_REQUIRE(d->box && d->elem && d->mol && d->group_offsets && d->group_atoms, "NULL array in descriptor");
    frmc_engine *e = new frmc_engine();
    int rc = FRMC_OK;
    hipError_t err = hipSuccess;
#define EFAIL(code, ...)            \
    do {                            \
        set_error(__VA_ARGS__);     \
        rc = code;                  \
        goto fail;                  \
    } while (0)
#define ECHK(x)                                                                     \
    do {                                                                            \
        err = (x);                                                                  \
        if (err != hipSuccess) EFAIL(FRMC_EHIP, "%s: %s", #x, hipGetErrorString(err)); \
    } while (0)
    {
        e->n = d->n;
        e->nel = d->nel;
        e->T = d->nel * (d->nel + 1) / 2;
        e->C = d->nconstraints;
        e->G = d->ngroups;
        e->isPBC = d->isPBC ? 1 : 0;
        e->kmax = d->kmax;
        for (int q = 0; q < 9; q++) e->B.b[q] = d->basis[q];
        for (int q = 0; q < 9; q++) e->recip[q] = d->recip[q];
        e->ortho = e->isPBC && d->basis[1] == 0.f && d->basis[2] == 0.f && d->basis[3] == 0.f && d->basis[5] == 0.f &&
                   d->basis[6] == 0.f && d->basis[7] == 0.f;
        if (d->stream) {
            e->st = (hipStream_t)d->stream;
        } else {
            ECHK(hipStreamCreate(&e->st));
            e->own_stream = true;
        }
        // validate groups on the host
        e->h_goff.assign(d->group_offsets, d->group_offsets + d->ngroups + 1);
        if (e->h_goff[0] != 0) EFAIL(FRMC_EINVAL, "group_offsets[0] must be 0");
        for (int32_t gi = 0; gi < d->ngroups; gi++) {
            int32_t g = e->h_goff[gi + 1] - e->h_goff[gi];
            if (g < 1 || g > kMaxG) EFAIL(FRMC_EINVAL, "group %d has %d atoms (1..%d supported)", gi, g, kMaxG);
            e->gmax = std::max(e->gmax, g);
        }
        e->total_gatoms = e->h_goff[d->ngroups];
        e->h_gatoms.assign(d->group_atoms, d->group_atoms + e->total_gatoms);
        for (int32_t a : e->h_gatoms)
            if (a < 0 || a >= d->n) EFAIL(FRMC_EINVAL, "group atom index %d out of range [0,%d)", a, d->n);
        for (int32_t i = 0; i < d->n; i++)
            if (d->elem[i] < 0 || d->elem[i] >= d->nel)
                EFAIL(FRMC_EINVAL, "elementIndex[%d]=%d out of range [0,%d)", i, d->elem[i], d->nel);
        ECHK(dupload(e, &e->xyz, d->box, 3 * (size_t)d->n));
        ECHK(dupload(e, &e->dbasis, d->basis, 9));
        ECHK(dupload(e, &e->mol, d->mol, (size_t)d->n));
        ECHK(dupload(e, &e->elem, d->elem, (size_t)d->n));
        ECHK(dupload(e, &e->goff, d->group_offsets, (size_t)d->ngroups + 1));
        ECHK(dupload(e, &e->gatoms, d->group_atoms, (size_t)e->total_gatoms));
        e->h_box.assign(d->box, d->box + 3 * (size_t)d->n);
        if (d->real) {
            e->h_real.assign(d->real, d->real + 3 * (size_t)d->n);
            e->have_real = true;
        }
        // constraints
        int32_t off = 0;
        std::vector<int32_t> sfc;
        for (int ci = 0; ci < e->C; ci++) {
            const frmc_constraint_desc &cd = d->constraints[ci];
            HostC &hc = e->c[ci];
            hc.d = cd;
            DevC &dc = hc.dev;
            memset(&dc, 0, sizeof(dc));
            if (cd.kind != FRMC_PDC && cd.kind != FRMC_SFC) EFAIL(FRMC_EINVAL, "constraint %d: unknown kind", ci);
            if (cd.histSize < 1) EFAIL(FRMC_EINVAL, "constraint %d: histSize must be >= 1", ci);
            if (!(cd.bin > 0.f)) EFAIL(FRMC_EINVAL, "constraint %d: bin must be > 0", ci);
            dc.kind = cd.kind;
            dc.H = cd.histSize;
            dc.M = cd.kind == FRMC_PDC ? cd.histSize : cd.nQ;
            if (dc.M < 1) EFAIL(FRMC_EINVAL, "constraint %d: no model points", ci);
            dc.g = make_grid(cd.minDistance, cd.maxDistance, cd.bin, cd.histSize);
            dc.hasW = cd.dataWeights != nullptr;
            dc.hasSF = cd.scaleFactor != 1.0f;
            dc.sf = cd.scaleFactor;
            if (!cd.pairWeights || !cd.pairDensity || !cd.shellVolumes || !cd.prefactor || !cd.experimental)
                EFAIL(FRMC_EINVAL, "constraint %d: NULL array", ci);
            if (cd.kind == FRMC_SFC && !cd.sqMatrix) EFAIL(FRMC_EINVAL, "constraint %d: SFC needs sqMatrix", ci);
            const float *wp, *Dp, *sv, *pf, *ex, *dw = nullptr, *sm = nullptr;
            ECHK(dupload(e, (float **)&wp, cd.pairWeights, (size_t)e->T));
            ECHK(dupload(e, (float **)&Dp, cd.pairDensity, (size_t)e->T));
            ECHK(dupload(e, (float **)&sv, cd.shellVolumes, (size_t)dc.H));
            ECHK(dupload(e, (float **)&pf, cd.prefactor, (size_t)dc.H));
            ECHK(dupload(e, (float **)&ex, cd.experimental, (size_t)dc.M));
            if (cd.dataWeights) ECHK(dupload(e, (float **)&dw, cd.dataWeights, (size_t)dc.M));
            if (cd.kind == FRMC_SFC) ECHK(dupload(e, (float **)&sm, cd.sqMatrix, (size_t)dc.H * dc.M));
            dc.w = wp; dc.D = Dp; dc.shellV = sv; dc.pref = pf; dc.expv = ex; dc.dataW = dw; dc.sqm = sm;
            std::vector<int32_t> lv, pg;
            pairwise_plan(dc.M, lv, pg);
            if (pg.size() > 4096) EFAIL(FRMC_EINVAL, "constraint %d: model too long", ci);
            const int32_t *lvp, *pgp;
            ECHK(dupload(e, (int32_t **)&lvp, lv.data(), lv.size()));
            ECHK(dupload(e, (int32_t **)&pgp, pg.data(), pg.size()));
            dc.leaves = lvp; dc.prog = pgp;
            dc.nleaves = (int32_t)(lv.size() / 2);
            dc.nprog = (int32_t)pg.size();
            ECHK(dalloc(e, &dc.base, (size_t)e->T * dc.H));
            ECHK(dalloc(e, &dc.slot, 2 * (size_t)e->nel * e->nel * dc.H + 1));
            dc.dnOff = off;
            off += e->T * dc.H;
            e->Hmax = std::max(e->Hmax, dc.H);
            e->Mmax = std::max(e->Mmax, dc.M);
            if (cd.kind == FRMC_SFC) sfc.push_back(ci);
        }
        e->dnStride = off;
        if ((size_t)e->dnStride * 4 > 150 * 1024)
            EFAIL(FRMC_EINVAL, "pair-type histograms of all constraints need %d KB of LDS (max 150)",
                  e->dnStride * 4 / 1024);
        e->nsfc = (int32_t)sfc.size();
        ECHK(dupload(e, &e->sfc_ids, sfc.data(), sfc.size()));
        // window buffers
        const size_t K = (size_t)e->kmax;
        ECHK(dalloc(e, &e->pgroup, K));
        ECHK(dalloc(e, &e->poff, K + 1));
        ECHK(dalloc(e, &e->pxyz, K * e->gmax * 3));
        ECHK(dalloc(e, &e->dn, K * e->dnStride));
        ECHK(dalloc(e, &e->Gb, K * e->C * e->Hmax));
        ECHK(dalloc(e, &e->sqv, K * e->C * e->Mmax));
        ECHK(dalloc(e, &e->chi2, K * e->C));
        ECHK(dalloc(e, &e->tse, K));
        ECHK(hipHostMalloc((void **)&e->h_pgroup, sizeof(int32_t) * K, hipHostMallocDefault));
        ECHK(hipHostMalloc((void **)&e->h_poff, sizeof(int32_t) * (K + 1), hipHostMallocDefault));
        ECHK(hipHostMalloc((void **)&e->h_pxyz, sizeof(float) * K * e->gmax * 3, hipHostMallocDefault));
        ECHK(hipHostMalloc((void **)&e->h_out, sizeof(float) * K * (1 + e->C), hipHostMallocDefault));
        ECHK(hipDeviceSynchronize());
    }
#undef ECHK
#undef EFAIL
    *out = e;
    return FRMC_OK;
fail:
    free_engine(e);
    return rc;
}

int frmc_engine_destroy(frmc_engine *e) {
    free_engine(e);
    return FRMC_OK;
}

int frmc_engine_compute_data(frmc_engine *e, float *chi2_out, float *tse_out) {
    FRMC_REQUIRE(e, "NULL engine");
    hipStream_t st = e->st;
    for (int ci = 0; ci < e->C; ci++) {
        DevC &dc = e->c[ci].dev;
        HistArgs A;
        A.g = dc.g;
        A.nel = e->nel;
        A.size = (int64_t)e->nel * e->nel * dc.H;
        k_fill_i32_e<<<1024, 256, 0, st>>>(dc.slot, 2 * A.size + 1, 0);
        FRMC_CHECK_LAUNCH();
        int rc = launch_full(e->xyz, e->n, e->dbasis, e->isPBC, e->ortho != 0, e->mol, e->elem, A, dc.slot, st);
        if (rc) return rc;
        k_symmetrise<<<blocks_for((int64_t)e->T * dc.H, 256), 256, 0, st>>>(dc.base, dc.slot, e->nel, dc.H);
        FRMC_CHECK_LAUNCH();
    }
    return frmc_engine_refresh(e, chi2_out, tse_out);
}

int frmc_engine_refresh(frmc_engine *e, float *chi2_out, float *tse_out) {
    // evaluate the current state as a zero-delta proposal
    FRMC_REQUIRE(e, "NULL engine");
    hipStream_t st = e->st;
    // proposal 0 = group 0 moved onto its own current coordinates (delta 0)
    e->h_pgroup[0] = 0;
    e->h_poff[0] = 0;
    int32_t g = e->h_goff[1] - e->h_goff[0];
    e->h_poff[1] = g;
    FRMC_CHECK_HIP(hipMemcpyAsync(e->pgroup, e->h_pgroup, sizeof(int32_t), hipMemcpyHostToDevice, st));
    FRMC_CHECK_HIP(hipMemcpyAsync(e->poff, e->h_poff, 2 * sizeof(int32_t), hipMemcpyHostToDevice, st));
    FRMC_CHECK_HIP(hipStreamSynchronize(st));
    View V = make_view(e);
    (void)V;
    k_fill_i32_e<<<64, 256, 0, st>>>(e->dn, e->dnStride, 0);
    FRMC_CHECK_LAUNCH();
    // k_gr / k_sq / k_finish on the zero delta
    {
        int32_t maxLeaves = 1;
        for (int i = 0; i < e->C; i++) maxLeaves = std::max(maxLeaves, e->c[i].dev.nleaves);
        size_t gshm = sizeof(float) * (e->Hmax + maxLeaves);
        if (gshm > 48 * 1024)
            (void)hipFuncSetAttribute((const void *)k_gr, hipFuncAttributeMaxDynamicSharedMemorySize, (int)gshm);
        k_gr<<<dim3(1, (unsigned)e->C), 256, gshm, st>>>(V);
        FRMC_CHECK_LAUNCH();
        if (e->nsfc > 0) {
            dim3 sg(blocks_for(e->Mmax, 256), 1, (unsigned)e->nsfc);
            k_sq<<<sg, 256, 0, st>>>(V, 1, e->sfc_ids);
            FRMC_CHECK_LAUNCH();
        }
        k_finish<<<1, 256, sizeof(float) * maxLeaves, st>>>(V);
        FRMC_CHECK_LAUNCH();
    }
    FRMC_CHECK_HIP(hipMemcpyAsync(e->h_out, e->tse, sizeof(float), hipMemcpyDeviceToHost, st));
    FRMC_CHECK_HIP(hipMemcpyAsync(e->h_out + 1, e->chi2, sizeof(float) * e->C, hipMemcpyDeviceToHost, st));
    FRMC_CHECK_HIP(hipStreamSynchronize(st));
    e->cur_tse = e->h_out[0];
    for (int ci = 0; ci < e->C; ci++) e->cur_chi2[ci] = e->h_out[1 + ci];
    if (tse_out) *tse_out = e->cur_tse;
    if (chi2_out)
        for (int ci = 0; ci < e->C; ci++) chi2_out[ci] = e->cur_chi2[ci];
    return FRMC_OK;
}

int frmc_engine_eval(frmc_engine *e, int32_t k, const int32_t *group_ids, const float *new_box, float *tse_out,
                     float *chi2_out) {
    FRMC_REQUIRE(e, "NULL engine");
    FRMC_REQUIRE(k >= 1 && k <= e->kmax, "k=%d out of range [1,%d]", k, e->kmax);
    FRMC_REQUIRE(group_ids && new_box && tse_out, "NULL argument");
    int32_t off = 0;
    for (int32_t p = 0; p < k; p++) {
        int32_t gi = group_ids[p];
        FRMC_REQUIRE(gi >= 0 && gi < e->G, "proposal %d: group %d out of range [0,%d)", p, gi, e->G);
        e->h_pgroup[p] = gi;
        e->h_poff[p] = off;
        off += e->h_goff[gi + 1] - e->h_goff[gi];
    }
    e->h_poff[k] = off;
    memcpy(e->h_pxyz, new_box, sizeof(float) * 3 * (size_t)off);
    hipStream_t st = e->st;
    FRMC_CHECK_HIP(hipMemcpyAsync(e->pgroup, e->h_pgroup, sizeof(int32_t) * k, hipMemcpyHostToDevice, st));
    FRMC_CHECK_HIP(hipMemcpyAsync(e->poff, e->h_poff, sizeof(int32_t) * (k + 1), hipMemcpyHostToDevice, st));
    FRMC_CHECK_HIP(hipMemcpyAsync(e->pxyz, e->h_pxyz, sizeof(float) * 3 * (size_t)off, hipMemcpyHostToDevice, st));
    int rc = launch_eval(e, k);
    if (rc) return rc;
    FRMC_CHECK_HIP(hipMemcpyAsync(e->h_out, e->tse, sizeof(float) * k, hipMemcpyDeviceToHost, st));
    FRMC_CHECK_HIP(hipMemcpyAsync(e->h_out + k, e->chi2, sizeof(float) * k * e->C, hipMemcpyDeviceToHost, st));
    FRMC_CHECK_HIP(hipStreamSynchronize(st));
    memcpy(tse_out, e->h_out, sizeof(float) * k);
    if (chi2_out) memcpy(chi2_out, e->h_out + k, sizeof(float) * k * e->C);
    e->last_k = k;
    return FRMC_OK;
}

int frmc_engine_commit(frmc_engine *e, int32_t p) {
    FRMC_REQUIRE(e, "NULL engine");
    FRMC_REQUIRE(p >= 0 && p < e->last_k, "proposal %d not in the last evaluated window (k=%d)", p, e->last_k);
    int rc = launch_commit(e, p);
    if (rc) return rc;
    e->cur_tse = e->h_out[p];
    for (int ci = 0; ci < e->C; ci++) e->cur_chi2[ci] = e->h_out[e->last_k + (int64_t)p * e->C + ci];
    // host mirrors of the box coordinates
    int32_t gi = e->h_pgroup[p];
    int32_t a0 = e->h_poff[p];
    for (int32_t a = 0; a < e->h_goff[gi + 1] - e->h_goff[gi]; a++) {
        int32_t id = e->h_gatoms[e->h_goff[gi] + a];
        for (int d = 0; d < 3; d++) e->h_box[3 * (size_t)id + d] = e->h_pxyz[3 * (size_t)(a0 + a) + d];
    }
    e->last_k = 0;  // the window buffers now describe a stale state
    return FRMC_OK;
}

int frmc_engine_sync(frmc_engine *e) {
    FRMC_REQUIRE(e, "NULL engine");
    FRMC_CHECK_HIP(hipStreamSynchronize(e->st));
    return FRMC_OK;
}

int frmc_engine_get_state(frmc_engine *e, float *box_out, int32_t constraint, int32_t *slot_out,
                          int32_t *base_out, float *chi2_out, float *tse_out) {
    FRMC_REQUIRE(e, "NULL engine");
    hipStream_t st = e->st;
    if (box_out)
        FRMC_CHECK_HIP(hipMemcpyAsync(box_out, e->xyz, sizeof(float) * 3 * (size_t)e->n, hipMemcpyDeviceToHost, st));
    if (slot_out || base_out) {
        FRMC_REQUIRE(constraint >= 0 && constraint < e->C, "constraint %d out of range", constraint);
        DevC &dc = e->c[constraint].dev;
        if (slot_out)
            FRMC_CHECK_HIP(hipMemcpyAsync(slot_out, dc.slot, sizeof(int32_t) * (2 * (size_t)e->nel * e->nel * dc.H + 1),
                                          hipMemcpyDeviceToHost, st));
        if (base_out)
            FRMC_CHECK_HIP(hipMemcpyAsync(base_out, dc.base, sizeof(int32_t) * (size_t)e->T * dc.H,
                                          hipMemcpyDeviceToHost, st));
    }
    FRMC_CHECK_HIP(hipStreamSynchronize(st));
    if (chi2_out)
        for (int ci = 0; ci < e->C; ci++) chi2_out[ci] = e->cur_chi2[ci];
    if (tse_out) *tse_out = e->cur_tse;
    return FRMC_OK;
}

}  // extern "C"

// ------------------------------------------------------ native move driver

namespace frmc {

// MT19937 (Matsumoto & Nishimura 1998), the generator behind CPython's
// random and NumPy's legacy RandomState; state layout as those expose it.
// `used` counts the 32-bit words drawn, so a stream can be rewound to any
// proposal boundary of a window by replaying that many words.
struct Stream {
    frmc_mt19937 *s;
    int32_t used;
    inline uint32_t next() {
        if (s->pos >= 624) {
            for (int i = 0; i < 624; i++) {
                uint32_t y = (s->mt[i] & 0x80000000u) | (s->mt[(i + 1) % 624] & 0x7fffffffu);
                s->mt[i] = s->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
            }
            s->pos = 0;
        }
        uint32_t y = s->mt[s->pos++];
        y ^= (y >> 11);
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= (y >> 18);
        used++;
        return y;
    }
    // random.random() and RandomState.random_sample(): 53-bit double
    inline double uniform() {
        uint32_t a = next() >> 5, b = next() >> 6;
        return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0);
    }
    // random.randint(0, n-1) == randrange(0, n) -> _randbelow_with_getrandbits(n)
    inline int32_t randbelow(int32_t n) {
        int k = 0;
        while (k < 32 && ((uint32_t)1 << k) <= (uint32_t)n) k++;  // n.bit_length()
        uint32_t r;
        do {
            r = next() >> (32 - k);
        } while (r >= (uint32_t)n);
        return (int32_t)r;
    }
};

// np.linalg.norm of a float32 3-vector: x.dot(x) goes through BLAS sdot,
// which rounds each product to float32 and accumulates in double; NumPy
// casts the sum to float32 and takes the float32 sqrt.
static inline float np_norm3(const float v[3]) {
    float p0 = v[0] * v[0], p1 = v[1] * v[1], p2 = v[2] * v[2];
    double acc = 0.0;
    acc += (double)p0;
    acc += (double)p1;
    acc += (double)p2;
    return sqrtf((float)acc);
}

// generate_random_vector(minAmp, maxAmp) (Core/Collection.py:674-701)
static inline void random_vector(Stream &py, Stream &np_, float minAmp, float maxAmp, float v[3]) {
    float norm;
    do {
        for (int d = 0; d < 3; d++) v[d] = (float)(1.0 - 2.0 * np_.uniform());
        norm = np_norm3(v);
    } while (norm == 0.0f);
    for (int d = 0; d < 3; d++) v[d] = v[d] / norm;
    float base[3];
    for (int d = 0; d < 3; d++) base[d] = v[d] * minAmp;
    const float amp = maxAmp - minAmp;
    const float scale = (float)py.uniform() * amp;  // Python float * float32 -> float32
    for (int d = 0; d < 3; d++) v[d] = v[d] * scale;
    for (int d = 0; d < 3; d++) v[d] = v[d] + base[d];
}

// transform_coordinates (boundary_conditions_collection.pyx:105-108)
static inline void to_box(const float M[9], const float r[3], float b[3]) {
    b[0] = r[0] * M[0] + r[1] * M[3] + r[2] * M[6];
    b[1] = r[0] * M[1] + r[1] * M[4] + r[2] * M[7];
    b[2] = r[0] * M[2] + r[1] * M[5] + r[2] * M[8];
}

}  // namespace frmc

extern "C" int frmc_engine_run_translation(frmc_engine *e, int64_t nsteps, int32_t k, float minAmp, float maxAmp,
                                           float tolerance, frmc_mt19937 *py_state, frmc_mt19937 *np_state,
                                           frmc_run_stats *stats, int32_t *group_log, uint8_t *accept_log,
                                           float *tse_log) {
    FRMC_REQUIRE(e && py_state && np_state && stats, "NULL argument");
    FRMC_REQUIRE(e->have_real, "engine was created without real coordinates");
    FRMC_REQUIRE(k >= 1 && k <= e->kmax, "k=%d out of range [1,%d]", k, e->kmax);
    FRMC_REQUIRE(py_state->pos >= 0 && py_state->pos <= 624 && np_state->pos >= 0 && np_state->pos <= 624,
                 "bad MT19937 position");
    std::vector<int32_t> gids(k), aoff(k + 1);
    std::vector<float> nbox((size_t)k * e->gmax * 3), nreal((size_t)k * e->gmax * 3);
    std::vector<double> uacc(k);
    std::vector<int32_t> py_pre(k), py_post(k), np_used(k);
    std::vector<float> tse(k);
    int64_t step = 0;
    while (step < nsteps) {
        const int32_t kw = (int32_t)std::min<int64_t>(k, nsteps - step);
        const frmc_mt19937 py0 = *py_state, np0 = *np_state;
        Stream py{py_state, 0}, np_{np_state, 0};
        int32_t off = 0;
        for (int32_t p = 0; p < kw; p++) {
            // Engine.__on_runtime_step_select_group -> RandomSelector.select_index
            const int32_t gi = py.randbelow(e->G);
            gids[p] = gi;
            // TranslationGenerator.transform_coordinates
            float v[3];
            random_vector(py, np_, minAmp, maxAmp, v);
            py_pre[p] = py.used;
            np_used[p] = np_.used;
            uacc[p] = py.uniform();  // the acceptance draw, used only if the move is worse
            py_post[p] = py.used;
            aoff[p] = off;
            const int32_t g0 = e->h_goff[gi], g = e->h_goff[gi + 1] - g0;
            for (int32_t a = 0; a < g; a++) {
                const int32_t id = e->h_gatoms[g0 + a];
                float r[3], b[3];
                for (int d = 0; d < 3; d++) r[d] = e->h_real[3 * (size_t)id + d] + v[d];
                to_box(e->recip, r, b);
                for (int d = 0; d < 3; d++) {
                    nreal[3 * (size_t)(off + a) + d] = r[d];
                    nbox[3 * (size_t)(off + a) + d] = b[d];
                }
            }
            off += g;
        }
        aoff[kw] = off;
        int rc = frmc_engine_eval(e, kw, gids.data(), nbox.data(), tse.data(), nullptr);
        if (rc) return rc;
        stats->windows++;
        stats->evaluated += kw;
        // walk the window in order (Engine.py:3308-3338)
        int32_t acc_p = -1;
        bool consumed_u = false;
        for (int32_t p = 0; p < kw; p++) {
            stats->generated++;
            stats->tried++;
            bool accept = true;
            consumed_u = false;
            if (tse[p] > e->cur_tse) {
                consumed_u = true;
                // random() > tolerance, with the Python float compared as float32
                if ((float)uacc[p] > tolerance) accept = false;
                else stats->tolerated++;
            }
            if (group_log) group_log[step] = gids[p];
            if (accept_log) accept_log[step] = accept ? 1 : 0;
            if (tse_log) tse_log[step] = accept ? tse[p] : e->cur_tse;
            step++;
            if (accept) {
                stats->accepted++;
                acc_p = p;
                break;
            }
        }
        if (acc_p >= 0) {
            // rewind both streams to just after the accepted proposal's draws
            *py_state = py0;
            *np_state = np0;
            Stream rp{py_state, 0}, rn{np_state, 0};
            const int32_t npy = consumed_u ? py_post[acc_p] : py_pre[acc_p];
            while (rp.used < npy) (void)rp.next();
            while (rn.used < np_used[acc_p]) (void)rn.next();
            rc = frmc_engine_commit(e, acc_p);
            if (rc) return rc;
            const int32_t gi = gids[acc_p], g0 = e->h_goff[gi], g = e->h_goff[gi + 1] - g0;
            for (int32_t a = 0; a < g; a++) {
                const int32_t id = e->h_gatoms[g0 + a];
                for (int d = 0; d < 3; d++) e->h_real[3 * (size_t)id + d] = nreal[3 * (size_t)(aoff[acc_p] + a) + d];
            }
        }
        // without an acceptance both streams already stand after the window
    }
    FRMC_CHECK_HIP(hipStreamSynchronize(e->st));
    return FRMC_OK;
}

extern "C" int frmc_engine_set_counts(frmc_engine *e, int32_t constraint, const int32_t *slot_in) {
    FRMC_REQUIRE(e && slot_in, "NULL argument");
    FRMC_REQUIRE(constraint >= 0 && constraint < e->C, "constraint %d out of range", constraint);
    DevC &dc = e->c[constraint].dev;
    const size_t count = 2 * (size_t)e->nel * e->nel * dc.H + 1;
    FRMC_CHECK_HIP(hipMemcpyAsync(dc.slot, slot_in, sizeof(int32_t) * count, hipMemcpyHostToDevice, e->st));
    k_symmetrise<<<blocks_for((int64_t)e->T * dc.H, 256), 256, 0, e->st>>>(dc.base, dc.slot, e->nel, dc.H);
    FRMC_CHECK_LAUNCH();
    FRMC_CHECK_HIP(hipStreamSynchronize(e->st));
    return FRMC_OK;
}

extern "C" int frmc_engine_set_real(frmc_engine *e, const float *real_in) {
    FRMC_REQUIRE(e && real_in, "NULL argument");
    e->h_real.assign(real_in, real_in + 3 * (size_t)e->n);
    e->have_real = true;
    return FRMC_OK;
}

extern "C" int frmc_engine_get_real(frmc_engine *e, float *real_out) {
    FRMC_REQUIRE(e && real_out, "NULL argument");
    FRMC_REQUIRE(e->have_real, "engine was created without real coordinates");
    memcpy(real_out, e->h_real.data(), sizeof(float) * 3 * (size_t)e->n);
    return FRMC_OK;
}
