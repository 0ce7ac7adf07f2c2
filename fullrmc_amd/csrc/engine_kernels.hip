// engine_kernels.hip -- the kernels of one speculative window and the commit
// of an accepted proposal.
//
//   k_delta   "moved group vs all N" before/after binning of k proposals for
//             every constraint grid in one pass (the distance is computed once
//             and binned per grid; the reference recomputes it per
//             constraint).  Grid (chunks of the N atoms, proposals); each
//             block accumulates int32 pair-type deltas in LDS and merges them
//             once (plain stores for one chunk per proposal, atomics into the
//             zeroed delta buffer otherwise).
//   k_eval_a  grid (Q chunks of 64, groups of 4 proposals), one wave per
//             proposal: base + delta -> G(r) in the reference's float32
//             operation order (PairDistributionConstraints.py:854-895), the
//             PDC chi-square by NumPy's pairwise sum (:810-838) in the first
//             Q chunk, and S(Q) as the sequential-in-r float32 sum of
//             G[r]*M[r,q] (StructureFactorConstraints.py:772-773): one Q per
//             lane, M staged through LDS in tiles of 32 rows shared by the 4
//             proposals of the block.
//   k_eval_b  grid k: SFC chi-squares (pairwise), the total np.sum([chi2_c])
//             (Engine.py:3026-3029), results and a completion flag written to
//             pinned host memory; also zeroes the other delta buffer.
//   k_commit  data - before + after for the accepted proposal: slot-exact
//             counts (M - F slot convention, PairDistributionConstraints.py:
//             1053-1078) and pair-type sums by atomics, then the new
//             coordinates of the group (written only by block 0; no other
//             block reads the group's coordinates).
#include <algorithm>

#include "engine.h"

namespace frmc {

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

// Flat slot index [ea][eb][bin] of an in-range pair (bin == H aliases the
// next slot as in the reference); -1 past the (nel,nel,H) array.
__device__ __forceinline__ int64_t slot_index(float s, const DevC &c, int32_t ea, int32_t eb, int32_t nel) {
    int64_t flat = ((int64_t)ea * nel + eb) * c.H + bin_of(s, c.g);
    return (flat < 0 || flat >= (int64_t)nel * nel * c.H) ? -1 : flat;
}

// pair-type index (t*H + bin) of a flat slot index
__device__ __forceinline__ int32_t type_of_slot(int64_t flat, const DevC &c, int32_t nel) {
    int32_t a = (int32_t)(flat / ((int64_t)nel * c.H));
    int32_t b = (int32_t)((flat / c.H) % nel);
    return pair_type(a, b, nel) * c.H + (int32_t)(flat % c.H);
}

// group-internal pair q of a group of g atoms -> positions a > b
__device__ __forceinline__ void decode_pair(int32_t q, int32_t &a, int32_t &b) {
    a = 1;
    while (q >= a) {
        q -= a;
        a++;
    }
    b = q;
}

struct GroupLds {
    float o[kMaxG][3], w[kMaxG][3];
    int32_t id[kMaxG], el[kMaxG], mol[kMaxG];
};

__device__ __forceinline__ int32_t load_group(const View &V, int32_t p, GroupLds &G) {
    const int32_t gi = V.pgroup[p];
    const int32_t a0 = V.poff[p];
    const int32_t g = V.poff[p + 1] - a0;
    const int32_t *atoms = V.gatoms + V.goff[gi];
    for (int32_t a = threadIdx.x; a < g; a += blockDim.x) {
        const int32_t id = atoms[a];
        G.id[a] = id;
        G.el[a] = V.elem[id];
        G.mol[a] = V.mol[id];
        for (int d = 0; d < 3; d++) {
            G.o[a][d] = V.pold[3 * (int64_t)(a0 + a) + d];
            G.w[a][d] = V.pnew[3 * (int64_t)(a0 + a) + d];
        }
    }
    return g;
}

// ----------------------------------------------------------------- k_delta

template <bool PBC, bool ORTHO>
__global__ void __launch_bounds__(256) k_delta(View V, Inline I, int32_t chunkLen, int32_t merge_atomic) {
    extern __shared__ int32_t lds[];  // V.dnStride ints
    __shared__ GroupLds G;
    const int32_t p = blockIdx.y;
    const int tid = threadIdx.x;
    for (int32_t s = tid; s < V.dnStride; s += blockDim.x) lds[s] = 0;
    int32_t g;
    if (I.k > 0) {
        g = 1;
        if (tid == 0) {
            const int32_t id = I.atom[p];
            G.id[0] = id;
            G.el[0] = V.elem[id];
            G.mol[0] = V.mol[id];
            for (int d = 0; d < 3; d++) {
                G.o[0][d] = V.xyz[3 * (int64_t)id + d];
                G.w[0][d] = I.nw[p][d];
            }
        }
    } else {
        g = load_group(V, p, G);
    }
    __syncthreads();

    const int32_t j0 = blockIdx.x * chunkLen;
    const int32_t j1 = min(V.n, j0 + chunkLen);
    for (int32_t j = j0 + tid; j < j1; j += blockDim.x) {
        bool member = false;
        for (int32_t a = 0; a < g; a++) member |= (G.id[a] == j);
        if (member) continue;
        const float cx = V.xyz[3 * (int64_t)j], cy = V.xyz[3 * (int64_t)j + 1], cz = V.xyz[3 * (int64_t)j + 2];
        const int32_t ej = V.elem[j];
        for (int32_t a = 0; a < g; a++) {
            // M counts (atom a, other j) with diff = c[a] - c[j], before and after the move
            const float so = eng_pair_s<PBC, ORTHO>(G.o[a][0], G.o[a][1], G.o[a][2], cx, cy, cz, V.B);
            const float sn = eng_pair_s<PBC, ORTHO>(G.w[a][0], G.w[a][1], G.w[a][2], cx, cy, cz, V.B);
            for (int32_t ci = 0; ci < V.C; ci++) {
                const DevC &c = V.c[ci];
                if (in_range(so, c.g)) {
                    int64_t f = slot_index(so, c, G.el[a], ej, V.nel);
                    if (f >= 0) atomicSub(&lds[c.dnOff + type_of_slot(f, c, V.nel)], 1);
                }
                if (in_range(sn, c.g)) {
                    int64_t f = slot_index(sn, c, G.el[a], ej, V.nel);
                    if (f >= 0) atomicAdd(&lds[c.dnOff + type_of_slot(f, c, V.nel)], 1);
                }
            }
        }
    }
    // group-internal pairs: M - F leaves each unordered pair once, in slot
    // [e(later position), e(earlier position)]
    if (blockIdx.x == 0 && g > 1) {
        for (int32_t q = tid; q < g * (g - 1) / 2; q += blockDim.x) {
            int32_t a, b;
            decode_pair(q, a, b);
            const float so = eng_pair_s<PBC, ORTHO>(G.o[a][0], G.o[a][1], G.o[a][2], G.o[b][0], G.o[b][1], G.o[b][2], V.B);
            const float sn = eng_pair_s<PBC, ORTHO>(G.w[a][0], G.w[a][1], G.w[a][2], G.w[b][0], G.w[b][1], G.w[b][2], V.B);
            for (int32_t ci = 0; ci < V.C; ci++) {
                const DevC &c = V.c[ci];
                if (in_range(so, c.g)) {
                    int64_t f = slot_index(so, c, G.el[a], G.el[b], V.nel);
                    if (f >= 0) atomicSub(&lds[c.dnOff + type_of_slot(f, c, V.nel)], 1);
                }
                if (in_range(sn, c.g)) {
                    int64_t f = slot_index(sn, c, G.el[a], G.el[b], V.nel);
                    if (f >= 0) atomicAdd(&lds[c.dnOff + type_of_slot(f, c, V.nel)], 1);
                }
            }
        }
    }
    __syncthreads();
    int32_t *out = V.dn + (int64_t)p * V.dnStride;
    if (merge_atomic) {
        for (int32_t s = tid; s < V.dnStride; s += blockDim.x) {
            const int32_t v = lds[s];
            if (v) atomicAdd(&out[s], v);
        }
    } else {
        for (int32_t s = tid; s < V.dnStride; s += blockDim.x) out[s] = lds[s];
    }
}

// ------------------------------------------------------ pairwise summation

// NumPy's pairwise float32 sum (np.add.reduce over a contiguous 1-D array):
// leaves of <= 128 elements use 8 strided accumulators combined
// ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)) and then the tail; leaves of < 8 are
// summed sequentially from 0.  Leaves and their combine tree follow NumPy's
// split n2 = n/2 rounded down to a multiple of 8 (host-built plan).
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

// Postfix combine of the leaf sums; the program and the value stack live in
// LDS (a private array would be placed in scratch memory, one dependent
// memory round trip per push or pop).
__device__ float run_prog(const int32_t *prog, int32_t nprog, const float *leafv, float *stk) {
    int sp = 0;
    for (int32_t k = 0; k < nprog; k++) {
        const int32_t op = prog[k];
        if (op >= 0) {
            stk[sp++] = leafv[op];
        } else {
            const float b = stk[--sp];
            const float x = stk[--sp];
            stk[sp++] = x + b;
        }
    }
    return stk[0];
}

constexpr int kStack = 40;  // postfix stack depth (tree depth of n <= 2^30)

// Pairwise sums of kPG arrays a[w*stride + 0..M) for the waves whose
// `active` flag is set; all threads of the block call it; wave w's lane 0
// receives its sum.  scratch: kPG*(nleaves + kStack) floats; lprog: nprog ints
// (a postfix program over L leaves has 2L-1 entries).
__device__ float block_pairwise_pg(const float *a, int32_t stride, const DevC &c, float *scratch, int32_t *lprog,
                                   bool active) {
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    for (int32_t k = threadIdx.x; k < c.nprog; k += blockDim.x) lprog[k] = c.prog[k];
    float *leafv = scratch + w * (c.nleaves + kStack);
    if (active)
        for (int32_t l = lane; l < c.nleaves; l += 64)
            leafv[l] = leaf_sum(a + w * stride + c.leaves[2 * l], c.leaves[2 * l + 1]);
    __syncthreads();
    float res = 0.0f;
    if (active && lane == 0) res = run_prog(lprog, c.nprog, leafv, leafv + c.nleaves);
    __syncthreads();
    return res;
}

// --------------------------------------------------------------- k_eval_a

__global__ void __launch_bounds__(256) k_eval_a(View V, int32_t k) {
    extern __shared__ float sm[];
    float *Gs = sm;                              // [kPG][Hmax]
    float *sq = Gs + kPG * V.Hmax;               // [kPG][Hmax] PDC squared deviations
    float *Mt = sq + kPG * V.Hmax;               // [2][kRT][kQC]
    float *scratch = Mt + 2 * kRT * kQC;         // [kPG][maxLeaves + kStack]
    int32_t *lprog = (int32_t *)(scratch + kPG * (V.maxLeaves + kStack));  // [2*maxLeaves]
    const int tid = threadIdx.x, w = tid / 64, lane = tid % 64;
    const int32_t p = blockIdx.y * kPG + w;
    const bool valid = p < k;
    const int32_t qc = blockIdx.x;
    const int32_t *dnp = V.dn + (int64_t)(valid ? p : 0) * V.dnStride;
    const int32_t T = V.T, Hmax = V.Hmax;
    for (int32_t ci = 0; ci < V.C; ci++) {
        // constraint fields in registers: keeps every load below unconditional
        // and out of per-element branches
        const DevC c = V.c[ci];
        const int32_t H = c.H, M = c.M;
        if (c.kind == FRMC_PDC && qc != 0) continue;  // uniform over the block
        const int32_t *dn = dnp + c.dnOff;
        const int32_t *base = c.base;
        if (valid) {
            // U bins per lane per pass, their loads issued together per pair type
            constexpr int U = 16;
            for (int32_t r0 = 0; r0 < H; r0 += 64 * U) {
                float Gr[U];
#pragma unroll
                for (int i = 0; i < U; i++) Gr[i] = 0.0f;
                for (int32_t t = 0; t < T; t++) {
                    int32_t bv[U], dv[U];
#pragma unroll
                    for (int i = 0; i < U; i++) {
                        const int32_t r = min(r0 + lane + 64 * i, H - 1);  // clamped: always in bounds
                        bv[i] = base[t * H + r];
                        dv[i] = dn[t * H + r];
                    }
                    const float wt = c.w[t], Dt = c.D[t];
#pragma unroll
                    for (int i = 0; i < U; i++) {
                        // Gr += wij*nij/Dij with nij integer-valued (exact in float32)
                        const float nij = (float)(bv[i] + dv[i]);
                        float term = wt * nij;
                        term = term / Dt;
                        Gr[i] = Gr[i] + term;
                    }
                }
#pragma unroll
                for (int i = 0; i < U; i++) {
                    const int32_t r = r0 + lane + 64 * i;
                    if (r >= H) continue;
                    float g = Gr[i] / c.shellV[r];           // Gr /= shellVolumes
                    float G = c.pref[r] * (g - 1.0f);         // (4*PI*r*rho0)*(Gr-1)
                    if (c.kind == FRMC_PDC) {
                        if (c.hasSF) G = G * c.sf;
                        const float diff = c.expv[r] - G;
                        float v = diff * diff;
                        if (c.hasW) v = c.dataW[r] * v;
                        sq[w * Hmax + r] = v;
                    } else {
                        Gs[w * Hmax + r] = G;
                    }
                }
            }
        }
        __syncthreads();
        if (c.kind == FRMC_PDC) {
            const float chi = block_pairwise_pg(sq, Hmax, c, scratch, lprog, valid);
            if (valid && lane == 0) V.chi2d[(int64_t)p * V.C + ci] = chi;
            continue;
        }
        // S(Q) for q = qc*kQC + lane: sequential float32 sum over r of G[r]*M[r,q].
        // M tiles of kRT rows x kQC columns are double-buffered in LDS; the next
        // tile is loaded into registers while the current one is consumed.
        const int32_t q = qc * kQC + lane;
        const bool qok = valid && q < M;
        const float *sqm = c.sqm;
        constexpr int PER = (kRT * kQC) / 256;
        float pre[PER];
        // element e of a tile: row e / kQC, column e % kQC; out-of-range rows
        // and columns read a clamped in-bounds address and store zero
        const int32_t col = min(qc * kQC + tid % kQC, M - 1);
        const bool colok = qc * kQC + tid % kQC < M;
        float acc = 0.0f;
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int32_t row = min(tid / kQC + (256 / kQC) * i, H - 1);
            const float v = sqm[(int64_t)row * M + col];
            pre[i] = (colok && tid / kQC + (256 / kQC) * i < H) ? v : 0.0f;
        }
#pragma unroll
        for (int i = 0; i < PER; i++) Mt[tid + 256 * i] = pre[i];
        __syncthreads();
        int buf = 0;
        for (int32_t r0 = 0; r0 < H; r0 += kRT) {
            const int32_t rows = min(kRT, H - r0);
            const bool more = r0 + kRT < H;
            if (more) {
#pragma unroll
                for (int i = 0; i < PER; i++) {
                    const int32_t rr = r0 + kRT + tid / kQC + (256 / kQC) * i;
                    const float v = sqm[(int64_t)min(rr, H - 1) * M + col];
                    pre[i] = (colok && rr < H) ? v : 0.0f;
                }
            }
            const float *Tl = Mt + buf * (kRT * kQC);
            if (qok) {
                const float *g = Gs + w * Hmax + r0;
                int32_t rr = 0;
                if (r0 == 0) {
                    acc = g[0] * Tl[lane];
                    rr = 1;
                }
                // groups of 16 rows: all LDS reads of a group are issued before
                // its (order-preserving) chain of adds
                for (; rr + 16 <= rows; rr += 16) {
                    float pr[16];
#pragma unroll
                    for (int i = 0; i < 16; i++) pr[i] = g[rr + i] * Tl[(rr + i) * kQC + lane];
#pragma unroll
                    for (int i = 0; i < 16; i++) acc = acc + pr[i];
                }
                for (; rr < rows; rr++) {
                    const float prod = g[rr] * Tl[rr * kQC + lane];
                    acc = acc + prod;
                }
            }
            if (more) {
                float *dst = Mt + (1 - buf) * (kRT * kQC);
#pragma unroll
                for (int i = 0; i < PER; i++) dst[tid + 256 * i] = pre[i];
            }
            __syncthreads();
            buf = 1 - buf;
        }
        if (qok) {
            float S = acc + 1.0f;
            if (c.hasSF) {
                float t = S - 1.0f;
                t = c.sf * t;
                S = t + 1.0f;
            }
            const float diff = c.expv[q] - S;
            float v = diff * diff;
            if (c.hasW) v = c.dataW[q] * v;
            V.sqbuf[((int64_t)p * V.C + ci) * V.Mmax + q] = v;
        }
    }
}

// --------------------------------------------------------------- k_eval_b

__global__ void __launch_bounds__(256) k_eval_b(View V, int32_t k) {
    extern __shared__ float sm[];  // Mmax + maxLeaves + kStack, then 2*maxLeaves ints
    float *scratch = sm + V.Mmax;
    int32_t *lprog = (int32_t *)(scratch + V.maxLeaves + kStack);
    const int32_t p = blockIdx.x;
    // zero a slice of the other delta buffer for the next window
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < V.zero_len;
         s += (int64_t)gridDim.x * blockDim.x)
        V.dn_zero[s] = 0;
    float tse = 0.0f;
    for (int32_t ci = 0; ci < V.C; ci++) {
        const DevC &c = V.c[ci];
        float chi;
        if (c.kind == FRMC_SFC) {
            const float *src = V.sqbuf + ((int64_t)p * V.C + ci) * V.Mmax;
            for (int32_t q = threadIdx.x; q < c.M; q += blockDim.x) sm[q] = src[q];
            for (int32_t q = threadIdx.x; q < c.nprog; q += blockDim.x) lprog[q] = c.prog[q];
            __syncthreads();
            for (int32_t l = threadIdx.x; l < c.nleaves; l += blockDim.x)
                scratch[l] = leaf_sum(sm + c.leaves[2 * l], c.leaves[2 * l + 1]);
            __syncthreads();
            chi = threadIdx.x == 0 ? run_prog(lprog, c.nprog, scratch, scratch + c.nleaves) : 0.0f;
            __syncthreads();
        } else {
            chi = V.chi2d[(int64_t)p * V.C + ci];
        }
        // np.sum([SD_c / varianceSquared]) with varianceSquared == 1: fewer
        // than 8 terms are summed sequentially from 0
        if (threadIdx.x == 0) {
            tse = tse + chi;
            V.out_chi2[(int64_t)p * V.C + ci] = chi;
        }
    }
    if (threadIdx.x == 0) {
        V.out_tse[p] = tse;
        __threadfence_system();
        V.out_done[p] = V.seq;
    }
}

// ---------------------------------------------------------------- k_commit

template <bool PBC, bool ORTHO>
__device__ __forceinline__ void commit_pair(const View &V, float s, int32_t ea, int32_t eb, bool intra, int32_t sign) {
    for (int32_t ci = 0; ci < V.C; ci++) {
        const DevC &c = V.c[ci];
        if (!in_range(s, c.g)) continue;
        const int64_t size = (int64_t)V.nel * V.nel * c.H;
        const int64_t f = slot_index(s, c, ea, eb, V.nel);
        if (f < 0) {
            if (sign > 0) atomicAdd(&c.slot[2 * size], 1);
            continue;
        }
        atomicAdd(&c.slot[(intra ? 0 : size) + f], sign);
        atomicAdd(&c.base[type_of_slot(f, c, V.nel)], sign);
    }
}

template <bool PBC, bool ORTHO>
__global__ void __launch_bounds__(256) k_commit(View V, int32_t p) {
    __shared__ GroupLds G;
    const int tid = threadIdx.x;
    const int32_t g = load_group(V, p, G);
    __syncthreads();
    const int32_t j = blockIdx.x * blockDim.x + tid;
    if (j < V.n) {
        bool member = false;
        for (int32_t a = 0; a < g; a++) member |= (G.id[a] == j);
        if (!member) {
            const float cx = V.xyz[3 * (int64_t)j], cy = V.xyz[3 * (int64_t)j + 1], cz = V.xyz[3 * (int64_t)j + 2];
            const int32_t ej = V.elem[j], mj = V.mol[j];
            for (int32_t a = 0; a < g; a++) {
                const float so = eng_pair_s<PBC, ORTHO>(G.o[a][0], G.o[a][1], G.o[a][2], cx, cy, cz, V.B);
                const float sn = eng_pair_s<PBC, ORTHO>(G.w[a][0], G.w[a][1], G.w[a][2], cx, cy, cz, V.B);
                const bool intra = G.mol[a] == mj;
                commit_pair<PBC, ORTHO>(V, so, G.el[a], ej, intra, -1);
                commit_pair<PBC, ORTHO>(V, sn, G.el[a], ej, intra, +1);
            }
        }
    }
    if (blockIdx.x == 0) {
        for (int32_t q = tid; q < g * (g - 1) / 2; q += blockDim.x) {
            int32_t a, b;
            decode_pair(q, a, b);
            const float so = eng_pair_s<PBC, ORTHO>(G.o[a][0], G.o[a][1], G.o[a][2], G.o[b][0], G.o[b][1], G.o[b][2], V.B);
            const float sn = eng_pair_s<PBC, ORTHO>(G.w[a][0], G.w[a][1], G.w[a][2], G.w[b][0], G.w[b][1], G.w[b][2], V.B);
            const bool intra = G.mol[a] == G.mol[b];
            commit_pair<PBC, ORTHO>(V, so, G.el[a], G.el[b], intra, -1);
            commit_pair<PBC, ORTHO>(V, sn, G.el[a], G.el[b], intra, +1);
        }
        for (int32_t a = tid; a < g; a += blockDim.x)
            for (int d = 0; d < 3; d++) V.xyz[3 * (int64_t)G.id[a] + d] = G.w[a][d];
    }
}

// base[t][r] = symmetrised slot sums (intra+inter, both orders)
__global__ void k_symmetrise(int32_t *__restrict__ base, const int32_t *__restrict__ slot, int32_t nel, int32_t H) {
    const int32_t T = nel * (nel + 1) / 2;
    const int64_t size = (int64_t)nel * nel * H;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < (int64_t)T * H;
         s += (int64_t)gridDim.x * blockDim.x) {
        const int32_t t = (int32_t)(s / H), r = (int32_t)(s % H);
        int32_t i = 0, rem = t;
        while (rem >= nel - i) {
            rem -= nel - i;
            i++;
        }
        const int32_t j = i + rem;
        const int64_t ij = ((int64_t)i * nel + j) * H + r, ji = ((int64_t)j * nel + i) * H + r;
        base[s] = (i == j) ? slot[ij] + slot[size + ij] : ((slot[ij] + slot[ji]) + slot[size + ij]) + slot[size + ji];
    }
}

__global__ void k_fill(int32_t *__restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// ------------------------------------------------------------------ launch

int launch_delta(const View &V, const Inline &I, int32_t k, int32_t nchunk, hipStream_t st) {
    if (nchunk < 1) nchunk = 1;
    int32_t chunkLen = (V.n + nchunk - 1) / nchunk;
    chunkLen = ((chunkLen + 255) / 256) * 256;
    nchunk = (V.n + chunkLen - 1) / chunkLen;
    const int merge_atomic = nchunk > 1;
    const size_t shm = sizeof(int32_t) * V.dnStride;
    const dim3 grid((unsigned)nchunk, (unsigned)k);
#define FRMC_DELTA(P, O)                                                                                     \
    do {                                                                                                     \
        if (shm > 48 * 1024)                                                                                 \
            (void)hipFuncSetAttribute((const void *)k_delta<P, O>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                      (int)shm);                                                             \
        k_delta<P, O><<<grid, 256, shm, st>>>(V, I, chunkLen, merge_atomic);                                  \
    } while (0)
    if (V.isPBC) {
        if (V.ortho) FRMC_DELTA(true, true); else FRMC_DELTA(true, false);
    } else {
        FRMC_DELTA(false, false);
    }
#undef FRMC_DELTA
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

size_t eval_lds_bytes(const View &V) {
    return sizeof(float) * ((size_t)2 * kPG * V.Hmax + 2 * kRT * kQC + (size_t)kPG * (V.maxLeaves + kStack) +
                            (size_t)2 * V.maxLeaves);
}

int launch_eval(const View &V, int32_t k, hipStream_t st) {
    const size_t shm_a = eval_lds_bytes(V);
    if (shm_a > 48 * 1024)
        (void)hipFuncSetAttribute((const void *)k_eval_a, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm_a);
    k_eval_a<<<dim3((unsigned)V.nqc, (unsigned)((k + kPG - 1) / kPG)), 256, shm_a, st>>>(V, k);
    FRMC_CHECK_LAUNCH();
    const size_t shm_b = sizeof(float) * (V.Mmax + V.maxLeaves + kStack + (size_t)2 * V.maxLeaves);
    if (shm_b > 48 * 1024)
        (void)hipFuncSetAttribute((const void *)k_eval_b, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm_b);
    k_eval_b<<<(unsigned)k, 256, shm_b, st>>>(V, k);
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

int launch_commit(const View &V, int32_t p, hipStream_t st) {
    const unsigned g = blocks_for(V.n, 256);
    if (V.isPBC) {
        if (V.ortho) k_commit<true, true><<<g, 256, 0, st>>>(V, p);
        else k_commit<true, false><<<g, 256, 0, st>>>(V, p);
    } else {
        k_commit<false, false><<<g, 256, 0, st>>>(V, p);
    }
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

int launch_symmetrise(int32_t *base, const int32_t *slot, int32_t nel, int32_t H, hipStream_t st) {
    const int32_t T = nel * (nel + 1) / 2;
    k_symmetrise<<<blocks_for((int64_t)T * H, 256), 256, 0, st>>>(base, slot, nel, H);
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

int launch_fill(int32_t *p, int64_t n, hipStream_t st) {
    unsigned g = blocks_for(n, 256);
    if (g > 2048) g = 2048;
    k_fill<<<g, 256, 0, st>>>(p, n);
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

// NumPy pairwise-sum leaves and combine program for length n.
static void plan_rec(int32_t start, int32_t len, std::vector<int32_t> &lv, std::vector<int32_t> &pg) {
    if (len <= 128) {
        pg.push_back((int32_t)(lv.size() / 2));
        lv.push_back(start);
        lv.push_back(len);
        return;
    }
    int32_t n2 = len / 2;
    n2 -= n2 % 8;
    plan_rec(start, n2, lv, pg);
    plan_rec(start + n2, len - n2, lv, pg);
    pg.push_back(-1);
}

void pairwise_plan(int32_t n, std::vector<int32_t> &leaves, std::vector<int32_t> &prog) {
    leaves.clear();
    prog.clear();
    plan_rec(0, n, leaves, prog);
}

}  // namespace frmc
