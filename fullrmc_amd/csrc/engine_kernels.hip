// engine_kernels.hip -- the kernels of one speculative window and the commit
// of an accepted proposal.
//
//   k_delta   "moved group vs all N" before/after binning of k proposals for
//             every constraint grid in one pass (the distance is computed once
//             and binned per grid; the reference recomputes it per
//             constraint).  Grid (chunks of 256 atoms, groups of kDP
//             proposals), one other atom per thread; int32 pair-type deltas
//             go by atomics into the proposal's zeroed block of the window's
//             delta buffer (a move touches a few thousand bins).  It first
//             commits the previous window's accepted single-atom move when
//             one is pending (the work of k_commit, folded in).
//   k_eval_a  grid (work units, proposals), one wave each: base + delta ->
//             G(r) in the reference's float32 operation order
//             (PairDistributionConstraints.py:854-895); per PDC constraint the
//             chi-square by NumPy's pairwise sum (:810-838), per 64 Q points of
//             an SFC constraint S(Q) as the sequential-in-r float32 sum of
//             G[r]*M[r,q] (StructureFactorConstraints.py:772-773), one Q per
//             lane, M rows read coalesced from HBM/L2.
//   k_eval_b  grid k: SFC chi-squares (pairwise), the total np.sum([chi2_c])
//             (Engine.py:3026-3029), results and a completion flag written to
//             pinned host memory; also zeroes the other delta buffer.
//   k_commit  data - before + after for the accepted proposal: slot-exact
//             counts (M - F slot convention, PairDistributionConstraints.py:
//             1053-1078) by atomics, pair-type sums += the proposal's deltas,
//             then the new coordinates of the group (written only by block 0;
//             the group's old and new coordinates come in the arguments).
#include <algorithm>

#include "engine.h"

namespace frmc {

__device__ __forceinline__ int32_t pair_type(int32_t a, int32_t b, int32_t nel) {
    int32_t lo = a < b ? a : b, hi = a < b ? b : a;
    return lo * (2 * nel - lo + 1) / 2 + (hi - lo);
}

// Flat index [ea][eb][b] of the C-ordered (nel,nel,H) slot array, -1 past its
// end: a bin b >= H (a distance just below maxDistance) aliases into a later
// slot exactly as the reference's unchecked write does.
__device__ __forceinline__ int32_t flat_slot(int32_t b, int32_t ea, int32_t eb, int32_t nel, int32_t H) {
    const int32_t f = (ea * nel + eb) * H + b;
    return f < nel * nel * H ? f : -1;
}

// Pair-type index t*H + r of an in-range pair (the symmetrised sums G(r)
// reads), consistent with flat_slot: direct for b < H, decoded from the flat
// slot in the rare aliased case; -1 past the array.
__device__ __forceinline__ int32_t type_index(int32_t b, int32_t ea, int32_t eb, int32_t nel, int32_t H) {
    if (b < H) return pair_type(ea, eb, nel) * H + b;
    const int32_t f = flat_slot(b, ea, eb, nel, H);
    if (f < 0) return -1;
    return pair_type(f / (nel * H), (f / H) % nel, nel) * H + f % H;
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
    int32_t g;
};

// proposal p of the staged window into G (all threads of the block call it)
__device__ __forceinline__ void load_group(const View &V, int32_t p, GroupLds &G) {
    const int32_t gi = load_dev(V.pgroup + p);
    const int32_t a0 = load_dev(V.poff + p);
    const int32_t g = load_dev(V.poff + p + 1) - a0;
    // the proposal's own atom list (moves of atoms that are not one group),
    // or its group's atoms
    const int32_t *atoms = V.patoms ? V.patoms + a0 : V.gatoms + V.goff[gi];
    if (threadIdx.x == 0) G.g = g;
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
}

// ----------------------------------------------------------------- k_delta

// -1 at the old and +1 at the new bin of one (group atom, other atom) pair,
// per distinct constraint grid (a constraint whose grid repeats the previous
// one shares its delta block)
template <bool PBC, bool ORTHO>
__device__ __forceinline__ void delta_pair(const View &V, int32_t *out, float so, float sn, int32_t ea, int32_t eb) {
    for (int32_t ci = 0; ci < V.C; ci++) {
        const DevC &c = V.c[ci];
        if (c.sameGrid) continue;
        const int32_t bo = in_range(so, c.g) ? bin_of(so, c.g) : -1;
        const int32_t bn = in_range(sn, c.g) ? bin_of(sn, c.g) : -1;
        if (bo == bn) continue;  // -1 and +1 in the same bin (or both out of range) cancel
        if (bo >= 0) {
            const int32_t t = type_index(bo, ea, eb, V.nel, c.H);
            if (t >= 0) atomicSub(&out[c.dnOff + t], 1);
        }
        if (bn >= 0) {
            const int32_t t = type_index(bn, ea, eb, V.nel, c.H);
            if (t >= 0) atomicAdd(&out[c.dnOff + t], 1);
        }
    }
}

// slot counts (intra/inter) of one pair at squared norm s, sign -1 or +1
template <bool PBC, bool ORTHO>
__device__ __forceinline__ void commit_pair(const View &V, float s, int32_t ea, int32_t eb, bool intra, int32_t sign) {
    int32_t b = -1;
    for (int32_t ci = 0; ci < V.C; ci++) {
        const DevC &c = V.c[ci];
        if (!c.sameGrid) b = in_range(s, c.g) ? bin_of(s, c.g) : -1;
        if (b < 0) continue;
        const int32_t size = V.nel * V.nel * c.H;
        const int32_t f = flat_slot(b, ea, eb, V.nel, c.H);
        if (f < 0) {
            if (sign > 0) atomicAdd(&c.slot[2 * size], 1);  // past the array: counted, not stored
            continue;
        }
        atomicAdd(&c.slot[(intra ? 0 : size) + f], sign);
    }
}

// grid (ceil(n/256), ceil(k/kDP)): one other atom per thread, kDP proposals
// per block; deltas go straight into the proposal's (zeroed) block of dn.
// P.valid: the previous window's accepted single-atom move is committed here
// first -- base += its delta block (grid-stride), its slot counts (blocks of
// the first proposal row), its new coordinates (one thread) -- and every
// thread reads the moved atom at its new position.
// Block (bx, by) of a (gx, gy) grid; dnw is the window's delta buffer; S
// gives the proposals of a single-atom window (S.inl()), else they come from
// the staging buffers through G.
template <bool PBC, bool ORTHO, class Src>
__device__ __forceinline__ void delta_body(const View &V, const Src &S, const Pending &P, int32_t *dnw, int32_t k,
                                           int32_t bx, int32_t by, int32_t gx, int32_t gy, GroupLds *G) {
    const int tid = threadIdx.x;
    const int32_t p0 = by * kDP;
    const int32_t np = min(kDP, k - p0);
    const bool inl = S.inl();
    if (!inl) {
        for (int32_t q = 0; q < np; q++) load_group(V, p0 + q, G[q]);
        __syncthreads();
    }
    const int32_t j = bx * blockDim.x + tid;
    const bool jok = j < V.n;
    const int32_t jj = min(j, V.n - 1);  // clamped: loads stay in bounds
    float cx = V.xyz[3 * (int64_t)jj], cy = V.xyz[3 * (int64_t)jj + 1], cz = V.xyz[3 * (int64_t)jj + 2];
    const int32_t ej = V.elem[jj];
    if (P.valid) {
        const int64_t gt = ((int64_t)by * gx + bx) * blockDim.x + tid;
        const int64_t gs = (int64_t)gx * gy * blockDim.x;
        for (int32_t ci = 0; ci < V.C; ci++) {
            const DevC &c = V.c[ci];
            for (int64_t s = gt; s < (int64_t)V.T * c.H; s += gs) c.base[s] += P.dn[c.dnOff + s];
        }
        if (by == 0 && jok && j != P.atom) {
            const float so = pair_s<PBC, ORTHO>(P.old[0], P.old[1], P.old[2], cx, cy, cz, V.B);
            const float sn = pair_s<PBC, ORTHO>(P.nw[0], P.nw[1], P.nw[2], cx, cy, cz, V.B);
            const bool intra = V.mol[P.atom] == V.mol[j];
            commit_pair<PBC, ORTHO>(V, so, P.el, ej, intra, -1);
            commit_pair<PBC, ORTHO>(V, sn, P.el, ej, intra, +1);
        }
        if (jj == P.atom) {
            cx = P.nw[0];
            cy = P.nw[1];
            cz = P.nw[2];
        }
        if (bx == 0 && by == 0 && tid == 0)
            for (int d = 0; d < 3; d++) V.xyz[3 * (int64_t)P.atom + d] = P.nw[d];
    }
    for (int32_t q = 0; q < np; q++) {
        const int32_t p = p0 + q;
        int32_t *out = dnw + (int64_t)p * V.dnStride;
        if (inl) {
            // single-atom proposal: everything about the moved atom comes with
            // the proposal (kernel arguments or the batch's proposal table)
            const int32_t id = S.atom(p);
            const int32_t ea = S.el(p);
            // M counts (moved atom a, other j) with diff = c[a] - c[j], before and after
            const float so = pair_s<PBC, ORTHO>(S.old(p, 0), S.old(p, 1), S.old(p, 2), cx, cy, cz, V.B);
            const float sn = pair_s<PBC, ORTHO>(S.nw(p, 0), S.nw(p, 1), S.nw(p, 2), cx, cy, cz, V.B);
            if (jok && j != id) delta_pair<PBC, ORTHO>(V, out, so, sn, ea, ej);
            continue;
        }
        const GroupLds &Gq = G[q];
        const int32_t g = Gq.g;
        bool member = false;
        for (int32_t a = 0; a < g; a++) member |= (Gq.id[a] == j);
        if (jok && !member) {
            for (int32_t a = 0; a < g; a++) {
                const float so = pair_s<PBC, ORTHO>(Gq.o[a][0], Gq.o[a][1], Gq.o[a][2], cx, cy, cz, V.B);
                const float sn = pair_s<PBC, ORTHO>(Gq.w[a][0], Gq.w[a][1], Gq.w[a][2], cx, cy, cz, V.B);
                delta_pair<PBC, ORTHO>(V, out, so, sn, Gq.el[a], ej);
            }
        }
        // group-internal pairs: M - F leaves each unordered pair once, in slot
        // [e(later position), e(earlier position)]
        if (bx == 0) {
            for (int32_t pq = tid; pq < g * (g - 1) / 2; pq += blockDim.x) {
                int32_t a, b;
                decode_pair(pq, a, b);
                const float so = pair_s<PBC, ORTHO>(Gq.o[a][0], Gq.o[a][1], Gq.o[a][2], Gq.o[b][0], Gq.o[b][1],
                                                    Gq.o[b][2], V.B);
                const float sn = pair_s<PBC, ORTHO>(Gq.w[a][0], Gq.w[a][1], Gq.w[a][2], Gq.w[b][0], Gq.w[b][1],
                                                    Gq.w[b][2], V.B);
                delta_pair<PBC, ORTHO>(V, out, so, sn, Gq.el[a], Gq.el[b]);
            }
        }
    }
}

// proposals of a single-atom window in the kernel arguments (k > 0)
struct InlineSrc {
    const Inline &I;
    __device__ bool inl() const { return I.k > 0; }
    __device__ int32_t atom(int32_t p) const { return I.atom[p]; }
    __device__ int32_t el(int32_t p) const { return I.el[p]; }
    __device__ float old(int32_t p, int d) const { return I.old[p][d]; }
    __device__ float nw(int32_t p, int d) const { return I.nw[p][d]; }
};

// proposals of a replica in the batch's proposal table
struct PropSrc {
    const Prop *pp;
    __device__ bool inl() const { return true; }
    __device__ int32_t atom(int32_t p) const { return pp[p].atom; }
    __device__ int32_t el(int32_t p) const { return pp[p].el; }
    __device__ float old(int32_t p, int d) const { return pp[p].old[d]; }
    __device__ float nw(int32_t p, int d) const { return pp[p].nw[d]; }
};

// The pending commit of a window: the kernel argument, or (device window
// loop) the one k_loop_step left in device memory.
// Its 12 words are read one per lane (vector loads, all in flight together)
// and broadcast to the wave.
__device__ __forceinline__ Pending pending_of(const View &V, const Pending &arg) {
    if (!V.pend_dev) return arg;
    constexpr int kWords = (int)(sizeof(Pending) / 4);
    const int lane = threadIdx.x & 63;
    const uint32_t w = lane < kWords ? ((const uint32_t *)V.pend_dev)[lane] : 0u;
    Pending P;
    uint32_t *dst = (uint32_t *)&P;
#pragma unroll
    for (int i = 0; i < kWords; i++) dst[i] = (uint32_t)__builtin_amdgcn_readlane((int)w, i);
    return P;
}

// grid (ceil(n/256), ceil(k/kDP)); PROPS: the proposals come from the
// device table V.props (device window loop)
template <bool PBC, bool ORTHO, bool PROPS>
__global__ void __launch_bounds__(256) k_delta(View V, Inline I, Pending Parg, int32_t k) {
    __shared__ GroupLds G[kDP];
    const Pending P = pending_of(V, Parg);
    if (PROPS) delta_body<PBC, ORTHO>(V, PropSrc{V.props}, P, V.dn, k, blockIdx.x, blockIdx.y, gridDim.x, gridDim.y, G);
    else delta_body<PBC, ORTHO>(V, InlineSrc{I}, P, V.dn, k, blockIdx.x, blockIdx.y, gridDim.x, gridDim.y, G);
}

// ------------------------------------------------------- k_delta, chunk walk

// proposals of a staged single-atom window (more proposals than the kernel
// arguments carry), read from the device copy of the staging buffer
struct StagedSrc {
    const View &V;
    __device__ int32_t atom(int32_t p) const { return V.gatoms[V.goff[V.pgroup[p]]]; }
    __device__ int32_t el(int32_t p) const { return V.elem[atom(p)]; }
    __device__ float old(int32_t p, int d) const { return V.pold[3 * (int64_t)V.poff[p] + d]; }
    __device__ float nw(int32_t p, int d) const { return V.pnew[3 * (int64_t)V.poff[p] + d]; }
};

// a box coordinate wrapped into [0, 1] (1 only by rounding, the same point
// as 0 on the unit circle)
__device__ __forceinline__ float wrap01(float x) { return x - floorf(x); }

// distance on the unit circle from f to the interval [lo, hi] (0 inside it)
__device__ __forceinline__ float circ_gap(float f, float lo, float hi) {
    if (f >= lo && f <= hi) return 0.0f;
    const float direct = f < lo ? lo - f : f - hi;
    const float around = f < lo ? (f + 1.0f) - hi : (lo + 1.0f) - f;
    return fminf(direct, around);
}

// True when every atom of a chunk whose wrapped fractional coordinates lie
// in the box bx (lo x,y,z, hi x,y,z) is farther than wcut from box position
// f in the reference's minimum image: the real-space difference r of a
// fractional difference b satisfies |r| >= |b_i| h_i for each axis i (h_i
// the spacing of the lattice planes across axis i), and |b_i| is at least
// the circular gap between f_i and [lo_i, hi_i].  wcut exceeds the largest
// binned distance by more than the float32 rounding of both sides.
__device__ __forceinline__ bool far_from(float fx, float fy, float fz, const float *bx, const View &V) {
    return circ_gap(wrap01(fx), bx[0], bx[3]) * V.hsp[0] > V.wcut ||
           circ_gap(wrap01(fy), bx[1], bx[4]) * V.hsp[1] > V.wcut ||
           circ_gap(wrap01(fz), bx[2], bx[5]) * V.hsp[2] > V.wcut;
}

struct NearProp {
    int32_t id, el, near;
    float o[3], w[3];
};

// Single-atom window of a periodic box over the spatially sorted copy: block
// (c, row) takes chunk c (one atom per thread) and proposals
// [row*pp, (row+1)*pp), pp <= kPruneP.  A proposal whose old and new positions
// are both beyond wcut of the chunk's box has no pair in range there and is
// skipped; the others run as in delta_body (same pairs, same bins; the
// atomics' order does not change integer sums).  The pending commit is
// applied as in delta_body: base += its deltas (grid-stride), its slot
// counts over the chunks near its old or new position (row 0), the moved
// atom read at its new position; the chunk holding it takes the new
// position into its box, and row 0 of that chunk writes the new position
// and box back (the other rows read either version and substitute the same
// values).  The chunk's atom, its box and the proposals load together at
// the start: one memory round trip before the pair work.
template <bool ORTHO, class Src>
__device__ __forceinline__ void delta_chunk_body(const View &V, const Src &S, const Pending &P, int32_t k,
                                                 int32_t pp) {
    __shared__ NearProp prop[kPruneP];
    const int tid = threadIdx.x;
    const int32_t c = blockIdx.x, row = blockIdx.y;
    const int32_t p0 = row * pp, np = min(pp, k - p0);
    const float4 a = V.sx[(int64_t)c * kChunkAtoms + tid];
    float bx[6];
#pragma unroll
    for (int d = 0; d < 6; d++) bx[d] = load_dev(V.cbox + 8 * (int64_t)c + d);
    NearProp q;
    if (tid < np) {
        const int32_t p = p0 + tid;
        q.id = S.atom(p);
        q.el = S.el(p);
        for (int d = 0; d < 3; d++) {
            q.o[d] = S.old(p, d);
            q.w[d] = S.nw(p, d);
        }
    }
    if (P.valid) {
        const int64_t gt = ((int64_t)row * gridDim.x + c) * blockDim.x + tid;
        const int64_t gs = (int64_t)gridDim.x * gridDim.y * blockDim.x;
        for (int32_t ci = 0; ci < V.C; ci++) {
            const DevC &dc = V.c[ci];
            for (int64_t s = gt; s < (int64_t)V.T * dc.H; s += gs) dc.base[s] += P.dn[dc.dnOff + s];
        }
    }
    const bool own = P.valid && P.slot / kChunkAtoms == c;
    if (own) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const float f = wrap01(P.nw[d]);
            bx[d] = fminf(bx[d], f);
            bx[3 + d] = fmaxf(bx[3 + d], f);
        }
        if (row == 0 && tid == 0) {
            V.sx[P.slot] = make_float4(P.nw[0], P.nw[1], P.nw[2], __int_as_float(P.atom * 64 + P.el));
            for (int d = 0; d < 6; d++) V.cbox[8 * (int64_t)c + d] = bx[d];
        }
    }
    if (P.valid && c == 0 && row == 0 && tid == 0)
        for (int d = 0; d < 3; d++) V.xyz[3 * (int64_t)P.atom + d] = P.nw[d];
    if (tid < np) {
        q.near = !(far_from(q.o[0], q.o[1], q.o[2], bx, V) && far_from(q.w[0], q.w[1], q.w[2], bx, V));
        prop[tid] = q;
    }
    const bool cm = row == 0 && P.valid &&
                    !(far_from(P.old[0], P.old[1], P.old[2], bx, V) && far_from(P.nw[0], P.nw[1], P.nw[2], bx, V));
    __syncthreads();
    const int32_t tag = __float_as_int(a.w);
    if (tag < 0) return;  // padding past the last atom
    const int32_t j = tag >> 6, ej = tag & 63;
    float cx = a.x, cy = a.y, cz = a.z;
    if (P.valid && j == P.atom) {
        cx = P.nw[0];
        cy = P.nw[1];
        cz = P.nw[2];
    }
    if (cm && j != P.atom) {
        const float so = pair_s<true, ORTHO>(P.old[0], P.old[1], P.old[2], cx, cy, cz, V.B);
        const float sn = pair_s<true, ORTHO>(P.nw[0], P.nw[1], P.nw[2], cx, cy, cz, V.B);
        const bool intra = V.mol[P.atom] == V.mol[j];
        commit_pair<true, ORTHO>(V, so, P.el, ej, intra, -1);
        commit_pair<true, ORTHO>(V, sn, P.el, ej, intra, +1);
    }
    for (int32_t i = 0; i < np; i++) {
        if (!prop[i].near || j == prop[i].id) continue;
        const NearProp &r = prop[i];
        const float so = pair_s<true, ORTHO>(r.o[0], r.o[1], r.o[2], cx, cy, cz, V.B);
        const float sn = pair_s<true, ORTHO>(r.w[0], r.w[1], r.w[2], cx, cy, cz, V.B);
        delta_pair<true, ORTHO>(V, V.dn + (int64_t)(p0 + i) * V.dnStride, so, sn, r.el, ej);
    }
}

// grid (nchunk, ceil(k/pp)); proposals in the kernel arguments (SRC 0), the
// staging buffer (1) or the device window loop's table (2)
template <bool ORTHO, int SRC>
__global__ void __launch_bounds__(256) k_delta_chunks(View V, Inline I, Pending Parg, int32_t k, int32_t pp) {
    const Pending P = pending_of(V, Parg);
    if (SRC == 0) delta_chunk_body<ORTHO>(V, InlineSrc{I}, P, k, pp);
    else if (SRC == 1) delta_chunk_body<ORTHO>(V, StagedSrc{V}, P, k, pp);
    else delta_chunk_body<ORTHO>(V, PropSrc{V.props}, P, k, pp);
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

// Pairwise sum of a[0..M) by a block: threads sum the leaves, thread 0 runs
// the combine program and returns the sum (all threads must call it).
// scratch: nleaves + kStack floats; lprog: nprog ints (a postfix program over
// L leaves has 2L-1 entries).
__device__ float block_pairwise(const float *a, const DevC &c, float *scratch, int32_t *lprog) {
    for (int32_t k = threadIdx.x; k < c.nprog; k += blockDim.x) lprog[k] = c.prog[k];
    for (int32_t l = threadIdx.x; l < c.nleaves; l += blockDim.x)
        scratch[l] = leaf_sum(a + c.leaves[2 * l], c.leaves[2 * l + 1]);
    __syncthreads();
    float res = 0.0f;
    if (threadIdx.x == 0) res = run_prog(lprog, c.nprog, scratch, scratch + c.nleaves);
    return res;
}

// --------------------------------------------------------------- k_eval_a

// One block of kEvalThreads threads per (work unit, proposal): grid (nunits,
// k).  A work unit is one r-space constraint (PDC, PCF), or 64 Q points of
// one S(Q) constraint (SFC, RSFC).  All threads rebuild the proposal's G(r)
// from base + delta in the reference's float32 operation order
// (PairDistributionConstraints.py:854-895) into LDS (a few bins each, so
// their loads go out in few rounds), then
//   PDC: squared deviations of G(r) and their NumPy pairwise sum (:810-838)
//        -> chi2d;
//   PCF: the same for g(r), the G(r) sum over the shell volumes, with the
//        scale factor applied through G(r) (PairCorrelationConstraints.py:
//        126-169);
//   SFC: S(Q) of its Q points, one per lane of wave 0, as the sequential float32 sum
//        over r of G[r]*M[r,q] (StructureFactorConstraints.py:772-778), the
//        matrix column streamed from HBM/L2 (transposed, a batch of rows
//        ahead in registers) and G broadcast from LDS -> squared deviations;
//   RSFC: the same sum without the +1, scaled without the shift
//        (StructureFactorConstraints.py:1255-1261).
// The S(Q) sum is a chain of H dependent adds per Q: the kernel is bound by
// that latency, so each wave carries one chain per lane and the grid spreads
// the (Q chunk, proposal) pairs over the CUs.
// A result for the host: the value and the window's sequence number in one
// 64-bit store to host memory, so the host needs no ordering between stores
// (it waits until every word carries the current window's number).
__device__ __forceinline__ void publish(unsigned long long *w, float v, uint32_t seq) {
    const unsigned long long word = ((unsigned long long)seq << 32) | (unsigned long long)__float_as_uint(v);
    __hip_atomic_store(w, word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// What k_eval_b does for a window whose engine has a single r-space
// constraint, done by k_eval_a instead (batched windows): publish the
// chi-square and the total, and zero the other delta buffer.
struct Fuse {
    int32_t *dn_zero;
    int64_t zero_len;
    uint32_t seq;
    int32_t nb;  // blocks of the window
};

// G(r) of proposal p (dn its delta block) into Gs[0, H), or for an r-space
// constraint the squared deviations of its model, by NT threads with U bins
// each per pass, in the reference's float32 order
// (PairDistributionConstraints.py:854-895).
template <int U, int NT>
__device__ __forceinline__ void build_g(const View &V, const DevC &c, const int32_t *dn, float *Gs) {
    const int tid = threadIdx.x;
    const int32_t H = c.H, T = V.T;
    const bool rsp = c.rspace;
    const int32_t *base = c.base;
    // U bins per lane per pass; every load is unconditional (clamped index,
    // valid pointer): a load under a divergent branch is followed by its
    // own wait, which serialises the batch
    constexpr int TMAX = 3;  // pair types (two elements) whose counts load in one round
    const float *expv = rsp ? c.expv : c.shellV;                 // S(Q): unused (expv has M entries)
    const float *dataW = (rsp && c.hasW) ? c.dataW : c.shellV;   // unused unless hasW
    for (int32_t r0 = 0; r0 < H; r0 += NT * U) {
        int32_t rr[U];
#pragma unroll
        for (int i = 0; i < U; i++) rr[i] = min(r0 + tid + NT * i, H - 1);
        float Gr[U];
        float sv[U], pf[U], ex[U], dw[U];
#pragma unroll
        for (int i = 0; i < U; i++) Gr[i] = 0.0f;
        if (T <= TMAX) {
            // every type's counts in one round of loads (types past T
            // reload the last one and are not summed)
            int32_t bv[TMAX][U], dv[TMAX][U];
#pragma unroll
            for (int t = 0; t < TMAX; t++) {
                const int32_t tc = min(t, T - 1);
#pragma unroll
                for (int i = 0; i < U; i++) {
                    bv[t][i] = base[tc * H + rr[i]];
                    dv[t][i] = dn[tc * H + rr[i]];
                }
            }
#pragma unroll
            for (int i = 0; i < U; i++) {
                sv[i] = c.shellV[rr[i]];
                pf[i] = c.pref[rr[i]];
                ex[i] = expv[rr[i]];
                dw[i] = dataW[rr[i]];
            }
            // pair types in order: Gr += wij*nij/Dij with nij integer-valued (exact in float32)
#pragma unroll
            for (int t = 0; t < TMAX; t++) {
                if (t >= T) break;
                const float wt = c.w[t], Dt = c.D[t];
#pragma unroll
                for (int i = 0; i < U; i++) {
                    const float nij = (float)(bv[t][i] + dv[t][i]);
                    float term = wt * nij;
                    term = term / Dt;
                    Gr[i] = Gr[i] + term;
                }
            }
        } else {
            int32_t bv[U], dv[U];
#pragma unroll
            for (int i = 0; i < U; i++) {
                bv[i] = base[rr[i]];
                dv[i] = dn[rr[i]];
            }
            // the per-bin constants load with the first counts
#pragma unroll
            for (int i = 0; i < U; i++) {
                sv[i] = c.shellV[rr[i]];
                pf[i] = c.pref[rr[i]];
                ex[i] = expv[rr[i]];
                dw[i] = dataW[rr[i]];
            }
            // pair types in order; the next type's counts load while this one is summed
            for (int32_t t = 0; t < T; t++) {
                const int32_t tn = min(t + 1, T - 1);
                int32_t nb[U], nd[U];
#pragma unroll
                for (int i = 0; i < U; i++) {
                    nb[i] = base[tn * H + rr[i]];
                    nd[i] = dn[tn * H + rr[i]];
                }
                const float wt = c.w[t], Dt = c.D[t];
#pragma unroll
                for (int i = 0; i < U; i++) {
                    // Gr += wij*nij/Dij with nij integer-valued (exact in float32)
                    const float nij = (float)(bv[i] + dv[i]);
                    float term = wt * nij;
                    term = term / Dt;
                    Gr[i] = Gr[i] + term;
                    bv[i] = nb[i];
                    dv[i] = nd[i];
                }
            }
        }
#pragma unroll
        for (int i = 0; i < U; i++) {
            float g = Gr[i] / sv[i];           // Gr /= shellVolumes
            float G = pf[i] * (g - 1.0f);      // (4*PI*r*rho0)*(Gr-1)
            float out = G;
            if (rsp) {
                float model;
                if (c.kind == FRMC_PCF) {
                    // g(r); a scale factor acts on G(r): g = 1 + (sf*G)/(4*PI*r*rho0)
                    model = g;
                    if (c.hasSF) {
                        const float Gsf = G * c.sf;
                        model = 1.0f + Gsf / pf[i];
                    }
                } else {
                    model = c.hasSF ? G * c.sf : G;
                }
                const float diff = ex[i] - model;
                float v = diff * diff;
                if (c.hasW) v = dw[i] * v;
                out = v;
            }
            if (r0 + tid + NT * i < H) Gs[r0 + tid + NT * i] = out;
        }
    }
}

// Work unit u of proposal p.  A unit is one r-space constraint (PDC, PCF) or
// one leaf of NumPy's pairwise sum over an S(Q) constraint's Q points (at
// most 128: waves 0 and 1 run the S(Q) sums of its first and second 64).
// Every thread builds the proposal's G(r) into LDS first.  dnw is the
// window's delta buffer, sm the block's dynamic LDS (eval_lds_bytes).
// V.leafmode (windows whose totals the host combines): an r-space unit
// publishes its chi-square, an S(Q) unit the sum of its leaf, and every
// block zeroes a slice of the other delta buffer (nblk blocks, b this one);
// otherwise the squared deviations go to sqbuf for k_eval_b.  F (batched
// windows of one r-space constraint): also k_eval_b's work for that
// constraint.
template <int U, bool IL>
__device__ __forceinline__ void eval_a_body(const View &V, const int32_t *dnw, int32_t p, int32_t u, float *sm,
                                            const Fuse *F, int64_t b, int64_t nblk) {
    const int32_t Hp = (V.Hmax + 15) & ~15;      // whole batches of 16 rows (float4 reads aligned)
    float *Gs = sm;                              // [Hp] G(r) or PDC squared deviations
    float *scratch = Gs + Hp;                    // [maxLeaves + kStack] (S(Q): the leaf's 128 deviations)
    int32_t *lprog = (int32_t *)(scratch + V.maxLeaves + kStack);  // [2*maxLeaves]
    const int tid = threadIdx.x, lane = tid % 64, w = tid / 64;
    // leaf layout: the other delta buffer is cleared for the next window
    // after this block's last barrier, so that no barrier waits for these
    // stores
    auto clear_next = [&] {
        if (V.leafmode)
            for (int64_t s = b * blockDim.x + tid; s < V.zero_len; s += nblk * blockDim.x) V.dn_zero[s] = 0;
    };
    int32_t ci = 0;
    while (ci + 1 < V.C && u >= V.c[ci + 1].uOff) ci++;
    // constraint fields in registers: keeps every load below unconditional
    const DevC c = V.c[ci];
    const int32_t H = c.H, M = c.M;
    const bool rsp = c.rspace;
    const int32_t *dn = dnw + (int64_t)p * V.dnStride + c.dnOff;
    // S(Q) unit: leaf l (Q points [q0, q0 + len), lane q = q0 + 64 w + lane,
    // waves 0 and 1), or 64 Q points from q0 (wave 0)
    int32_t leaf = 0, q0 = 0, len = 0;
    if (!rsp) {
        leaf = u - c.uOff;
        if (V.leafunits) {
            q0 = c.leaves[2 * leaf];
            len = c.leaves[2 * leaf + 1];
        } else {
            q0 = leaf * kQC;
            len = min(kQC, M - q0);
        }
    }
    const bool chain_wave = !rsp && w < 2 && 64 * w < len;
    const int32_t q = q0 + 64 * w + lane;
    const bool qok = chain_wave && 64 * w + lane < len;
    const int32_t qc = min(q, M - 1);
    const float *mcol = c.sqm + (int64_t)qc * c.hpb;  // transposed: M[r][q] at mcol[r], 16-byte aligned
    const float *mil = c.sqi + 4 * (int64_t)qc;         // interleaved: M[r][q] at mil[(r/4)*4M + r%4]
    // the S(Q) sum streams the matrix column in batches of B rows, NBM of
    // them in flight (16 x 4 measured 1.6x faster than 32 x 2 on 2,000 rows)
    constexpr int B = 16, NBM = 4;               // hpb is a multiple of 32 >= B
    float mr[NBM][B], gg[2][B];
    auto load_m = [&](float *m, int32_t r0) {
        // a batch past the end (a prefetch that is never used) rereads the last one
        const int32_t rc = min(r0, c.hpb - B);
#pragma unroll
        for (int i4 = 0; i4 < B / 4; i4++) {
            // IL: the wave's lanes read one contiguous kilobyte per four rows;
            // transposed: each lane reads its own column.  The interleaved
            // offset is 32-bit (hpb*M < 2^31, checked at engine creation): a
            // 64-bit multiply per load measured 7 us slower per launch on
            // 2,000 rows
            const float *src = IL ? mil + (uint32_t)((rc / 4 + i4) * 4 * M) : mcol + rc + 4 * i4;
            const float4 v = *reinterpret_cast<const float4 *>(src);
            m[4 * i4] = v.x; m[4 * i4 + 1] = v.y; m[4 * i4 + 2] = v.z; m[4 * i4 + 3] = v.w;
        }
    };
    // the first rows of the matrix column and the experimental value load
    // now, under the G(r) phase
    float ex_q = 0.0f, dw_q = 1.0f;
    if (chain_wave) {
#pragma unroll
        for (int bq = 0; bq < NBM - 1; bq++) load_m(mr[bq], bq * B);
        ex_q = c.expv[min(q, M - 1)];
        if (c.hasW) dw_q = c.dataW[min(q, M - 1)];
    }
    build_g<U, kEvalThreads>(V, c, dn, Gs);
    __syncthreads();
    if (rsp) {
        const float chi = block_pairwise(Gs, c, scratch, lprog);
        if (tid == 0) {
            V.chi2d[(int64_t)p * V.C + ci] = chi;
            if (V.leafmode) publish(V.out_res + (int64_t)p * V.resW + 1 + ci, chi, V.seq);
        }
        if (F) {
            // np.sum([chi]) from 0, as k_eval_b sums it
            if (tid == 0) {
                const float tse = 0.0f + chi;
                publish(V.out_res + (int64_t)p * V.resW + 1, chi, F->seq);
                V.tse_dev[p] = tse;
                publish(V.out_res + (int64_t)p * V.resW, tse, F->seq);
            }
            for (int64_t s = (int64_t)p * blockDim.x + tid; s < F->zero_len; s += (int64_t)F->nb * blockDim.x)
                F->dn_zero[s] = 0;
        }
        clear_next();
        return;
    }
    if (chain_wave) {
        // The S(Q) sum starts from +0, as NumPy's add reduction does (a sum of
        // -0 products is +0).
        float acc = 0.0f;
        // two register batches (matrix column and G rows) in turn: one batch loads
        // while the other's chain runs, which hides both the HBM/L2 and the LDS latency
        auto load_g = [&](float *gg, int32_t r0) {
            const float *gsrc = Gs + min(r0, Hp - B > 0 ? Hp - B : 0);
#pragma unroll
            for (int i4 = 0; i4 < B / 4; i4++) {
                const float4 gv = *reinterpret_cast<const float4 *>(gsrc + 4 * i4);
                gg[4 * i4] = gv.x; gg[4 * i4 + 1] = gv.y; gg[4 * i4 + 2] = gv.z; gg[4 * i4 + 3] = gv.w;
            }
        };
        // all products first: the adds then run back to back, each waiting
        // only on the previous add
        auto chain_full = [&](const float *m, const float *g) {
            float pr[B];
#pragma unroll
            for (int i = 0; i < B; i++) pr[i] = g[i] * m[i];
            __builtin_amdgcn_sched_barrier(0);  // keep the scheduler from interleaving the two loops
#pragma unroll
            for (int i = 0; i < B; i++) acc = acc + pr[i];
        };
        // a batch that may end past H: rows past H are not added
        auto chain = [&](const float *m, const float *g, int32_t r0) {
            float pr[B];
#pragma unroll
            for (int i = 0; i < B; i++) pr[i] = g[i] * m[i];
#pragma unroll
            for (int i = 0; i < B; i++)
                if (r0 + i < H) acc = acc + pr[i];
        };
        load_g(gg[0], 0);
        // whole rounds of NBM batches without branches, so that the loads of
        // the next NBM - 1 batches stay in flight across the adds; ring slot
        // j holds batch bi + j
        const int32_t nb = (H + B - 1) / B;  // batches, the last one possibly partial
        int32_t bi = 0;
        for (; bi + NBM <= nb - 1; bi += NBM) {
#pragma unroll
            for (int j = 0; j < NBM; j++) {
                load_m(mr[(j + NBM - 1) % NBM], (bi + j + NBM - 1) * B);
                load_g(gg[(j + 1) & 1], (bi + j + 1) * B);
                chain_full(mr[j], gg[j & 1]);
            }
        }
        // the last (at most NBM) batches: ring slots 0..NBM-2 hold batches
        // bi..bi+NBM-2; the one after them loads here
#pragma unroll
        for (int j = 0; j < NBM; j++) {
            if (bi + j < nb) {
                if (bi + j + NBM - 1 < nb) load_m(mr[(j + NBM - 1) % NBM], (bi + j + NBM - 1) * B);
                load_g(gg[(j + 1) & 1], (bi + j + 1) * B);
                chain(mr[j], gg[j & 1], (bi + j) * B);
            }
        }
        if (qok) {
            float S;
            if (c.kind == FRMC_RSFC) {
                S = c.hasSF ? c.sf * acc : acc;
            } else {
                S = acc + 1.0f;
                if (c.hasSF) {
                    float t = S - 1.0f;
                    t = c.sf * t;
                    S = t + 1.0f;
                }
            }
            const float diff = ex_q - S;
            float v = diff * diff;
            if (c.hasW) v = dw_q * v;
            if (V.leafmode) scratch[64 * w + lane] = v;
            else V.sqbuf[((int64_t)p * V.C + ci) * V.Mmax + q] = v;
        }
    }
    if (!V.leafmode) return;
    // the leaf's sum in NumPy's order (the host combines the leaves)
    __syncthreads();
    if (tid == 0) publish(V.out_res + (int64_t)p * V.resW + c.leafOff + leaf, leaf_sum(scratch, len), V.seq);
    clear_next();
}

// --------------------------------------------------------------- k_eval_b

// proposal p of a k-block grid: the SFC chi-squares and the total, published
// with the window's number seq; the blocks also zero zero_len ints of dn_zero
__device__ __forceinline__ void eval_b_body(const View &V, int32_t *dn_zero, int64_t zero_len, uint32_t seq,
                                            int32_t p, int32_t nb, float *sm) {
    float *scratch = sm + V.Mmax;
    int32_t *lprog = (int32_t *)(scratch + V.maxLeaves + kStack);
    // zero a slice of the other delta buffer for the next window
    for (int64_t s = (int64_t)p * blockDim.x + threadIdx.x; s < zero_len; s += (int64_t)nb * blockDim.x)
        dn_zero[s] = 0;
    // the r-space chi-squares (k_eval_a's partials) load first, under the
    // S(Q) sums
    float rchi[kMaxC];
#pragma unroll
    for (int ci = 0; ci < kMaxC; ci++)
        rchi[ci] = (ci < V.C && V.c[ci].rspace) ? load_dev(V.chi2d + (int64_t)p * V.C + ci) : 0.0f;
    float tse = 0.0f;
#pragma unroll
    for (int ci = 0; ci < kMaxC; ci++) {
        if (ci >= V.C) break;
        const DevC &c = V.c[ci];
        float chi;
        if (!c.rspace) {
            const float *src = V.sqbuf + ((int64_t)p * V.C + ci) * V.Mmax;
            for (int32_t q = threadIdx.x; q < c.M; q += blockDim.x) sm[q] = load_dev(src + q);
            for (int32_t q = threadIdx.x; q < c.nprog; q += blockDim.x) lprog[q] = c.prog[q];
            __syncthreads();
            for (int32_t l = threadIdx.x; l < c.nleaves; l += blockDim.x)
                scratch[l] = leaf_sum(sm + c.leaves[2 * l], c.leaves[2 * l + 1]);
            __syncthreads();
            chi = threadIdx.x == 0 ? run_prog(lprog, c.nprog, scratch, scratch + c.nleaves) : 0.0f;
            __syncthreads();
        } else {
            chi = rchi[ci];
        }
        // np.sum([SD_c / varianceSquared]) with varianceSquared == 1: fewer
        // than 8 terms are summed sequentially from 0
        if (threadIdx.x == 0) {
            tse = tse + chi;
            publish(V.out_res + (int64_t)p * V.resW + 1 + ci, chi, seq);
        }
    }
    if (threadIdx.x == 0) {
        V.tse_dev[p] = tse;
        publish(V.out_res + (int64_t)p * V.resW, tse, seq);
    }
}

// grid k (Mmax + maxLeaves + kStack floats, then 2*maxLeaves ints of LDS)
__global__ void __launch_bounds__(256) k_eval_b(View V, int32_t k) {
    extern __shared__ float sm[];
    eval_b_body(V, V.dn_zero, V.zero_len, V.seq, blockIdx.x, gridDim.x, sm);
}

// --------------------------------------------------------------- k_eval_a

// grid (nunits, k): block (u, p) evaluates unit u for proposal p; U bins
// per thread per G(r) pass.  XCD: a one-dimensional grid of
// ceil(nunits/8)*8 blocks per proposal, so that (workgroups going to the
// XCDs in turn) unit u always runs on XCD u mod 8 and each XCD's L2 holds
// the matrix columns of its units only; the padding blocks return at once.
// IL: the S(Q) matrix read in its interleaved copy (one contiguous kilobyte
// per wave and four rows: fewer cache lines per load than the transposed
// copy's 64, which two chain waves of a CU contend for, and no L2 thrash
// when many proposals stream it), else the transposed one.
template <int U, bool XCD, bool IL>
__global__ void __launch_bounds__(kEvalThreads) k_eval_a(View V, int32_t k) {
    extern __shared__ float sm[];
    int32_t u, p;
    int64_t b, nblk;
    if (XCD) {
        const int32_t nu8 = (V.nunits + 7) & ~7;
        u = blockIdx.x % nu8;
        p = blockIdx.x / nu8;
        b = (int64_t)p * V.nunits + u;
        nblk = (int64_t)V.nunits * k;
        if (u >= V.nunits) return;
    } else {
        u = blockIdx.x;
        p = blockIdx.y;
        b = (int64_t)p * gridDim.x + u;
        nblk = (int64_t)gridDim.x * gridDim.y;
    }
    eval_a_body<U, IL>(V, V.dn, p, u, sm, nullptr, b, nblk);
}

// ------------------------------------------------- k_eval_a, split chain

// The leaf layout's S(Q) sums with the products computed apart from the
// adds: in a block of kPcThreads threads, waves 0 and 1 (the chain waves,
// Q points [q0, q0 + 64) and [q0 + 64, q0 + len) of the leaf) only add,
// while the other fourteen waves load the interleaved matrix for both
// halves, multiply by G from LDS (one read per row group for both) and
// store the products to an
// LDS ring of two slots of kPcRows rows; one barrier per slot hands a slot
// over.  A chain wave then issues one LDS read per four rows and the adds,
// so its sum runs close to the latency of the dependent adds.  Every sum and
// product is the one eval_a_body computes, in the same order.
constexpr int kPcThreads = 1024;
constexpr int kPcRows = 128;               // rows per ring slot
constexpr int kPcGroups = kPcRows / 4;     // 4-row groups per slot
constexpr int kPcProd = 14;                // producer waves (each serves both chain waves)
constexpr int kPcPerProd = (kPcGroups + kPcProd - 1) / kPcProd;  // row groups per producer per slot
constexpr int kPcAhead = 2;                // slots of matrix rows a producer has in flight
constexpr int kPcRead = 8;                 // product groups a chain wave has in flight

// LDS of k_eval_pc: G(r) (Hp floats), the ring (2 slots x 2 halves x
// kPcGroups x 64 float4), then the leaf and pairwise scratch of eval_a_body
size_t pc_lds_bytes(const View &V) {
    const size_t hp = (size_t)((V.Hmax + 15) & ~15);
    const size_t ring = (size_t)2 * 2 * kPcGroups * 64 * 4;
    return sizeof(float) * (hp + ring + std::max<size_t>(V.maxLeaves + kStack, 128) + (size_t)2 * V.maxLeaves);
}

template <int U>
__device__ __forceinline__ void eval_pc_body(const View &V, int32_t p, int32_t u, float *sm, int64_t b, int64_t nblk) {
    const int32_t Hp = (V.Hmax + 15) & ~15;
    float *Gs = sm;
    float4 *ring = reinterpret_cast<float4 *>(sm + Hp);  // Hp is a multiple of 16: 16-byte aligned
    float *scratch = sm + Hp + 2 * 2 * kPcGroups * 64 * 4;
    int32_t *lprog = (int32_t *)(scratch + V.maxLeaves + kStack);
    const int tid = threadIdx.x, lane = tid % 64, w = tid / 64;
    // the other delta buffer is cleared for the next window after this
    // block's last barrier, so that no barrier waits for these stores
    auto clear_next = [&] {
        for (int64_t s = b * blockDim.x + tid; s < V.zero_len; s += nblk * blockDim.x) V.dn_zero[s] = 0;
    };
    int32_t ci = 0;
    while (ci + 1 < V.C && u >= V.c[ci + 1].uOff) ci++;
    const DevC c = V.c[ci];
    const int32_t H = c.H, M = c.M;
    const int32_t *dn = V.dn + (int64_t)p * V.dnStride + c.dnOff;
    if (c.rspace) {
        build_g<U, kPcThreads>(V, c, dn, Gs);
        __syncthreads();
        const float chi = block_pairwise(Gs, c, scratch, lprog);
        if (tid == 0) {
            V.chi2d[(int64_t)p * V.C + ci] = chi;
            publish(V.out_res + (int64_t)p * V.resW + 1 + ci, chi, V.seq);
        }
        clear_next();
        return;
    }
    const int32_t leaf = u - c.uOff;
    const int32_t q0 = c.leaves[2 * leaf], len = c.leaves[2 * leaf + 1];
    const bool chain = w < 2;
    const int32_t half = chain ? w : 0;                   // a chain wave's half of the leaf
    const int32_t pw = chain ? 0 : w - 2;                 // a producer's rank
    const int32_t q = q0 + 64 * half + lane;
    const int32_t qc = min(q, M - 1);
    const float *mil = c.sqi + 4 * (int64_t)qc;           // M[r][qc] at mil[(r/4)*4M + r%4]
    // a producer serves both halves: lane l computes Q points q0 + l and q0 + 64 + l
    const float *mil1 = c.sqi + 4 * (int64_t)min(q0 + 64 + lane, M - 1);
    const int32_t ngroups = c.hpb / 4;
    const int32_t nslot = (H + kPcRows - 1) / kPcRows;
    float4 mreg[kPcAhead + 1][2 * kPcPerProd];
    auto load_slot = [&](float4 *mr, int32_t s) {
#pragma unroll
        for (int j = 0; j < kPcPerProd; j++) {
            const int32_t rg = min(pw + kPcProd * j, kPcGroups - 1);
            const uint32_t off = (uint32_t)(min(s * kPcGroups + rg, ngroups - 1) * 4 * M);
            mr[2 * j] = *reinterpret_cast<const float4 *>(mil + off);
            mr[2 * j + 1] = *reinterpret_cast<const float4 *>(mil1 + off);
        }
    };
    if (!chain) {
#pragma unroll
        for (int s = 0; s < kPcAhead; s++) load_slot(mreg[s], s);
    }
    float ex_q = 0.0f, dw_q = 1.0f;
    if (chain) {
        ex_q = c.expv[qc];
        if (c.hasW) dw_q = c.dataW[qc];
    }
    build_g<U, kPcThreads>(V, c, dn, Gs);
    __syncthreads();
    if (chain) {
        // S(Q) from +0, as NumPy's add reduction does
        float acc = 0.0f;
        for (int32_t s = 0; s < nslot; s++) {
            __syncthreads();  // slot s written
            const float4 *src = ring + ((s & 1) * 2 + half) * kPcGroups * 64;
            float4 v[kPcRead];
#pragma unroll
            for (int i = 0; i < kPcRead; i++) v[i] = src[i * 64 + lane];
            const int32_t rows = min(kPcRows, H - s * kPcRows);
            if (rows == kPcRows) {
#pragma unroll
                for (int g = 0; g < kPcGroups; g++) {
                    const float4 x = v[g % kPcRead];
                    if (g + kPcRead < kPcGroups) v[g % kPcRead] = src[(g + kPcRead) * 64 + lane];
                    acc = acc + x.x;
                    acc = acc + x.y;
                    acc = acc + x.z;
                    acc = acc + x.w;
                }
            } else {
#pragma unroll
                for (int g = 0; g < kPcGroups; g++) {
                    const float4 x = v[g % kPcRead];
                    if (g + kPcRead < kPcGroups) v[g % kPcRead] = src[(g + kPcRead) * 64 + lane];
                    if (4 * g + 0 < rows) acc = acc + x.x;
                    if (4 * g + 1 < rows) acc = acc + x.y;
                    if (4 * g + 2 < rows) acc = acc + x.z;
                    if (4 * g + 3 < rows) acc = acc + x.w;
                }
            }
        }
        if (64 * half + lane < len) {
            float S;
            if (c.kind == FRMC_RSFC) {
                S = c.hasSF ? c.sf * acc : acc;
            } else {
                S = acc + 1.0f;
                if (c.hasSF) {
                    float t = S - 1.0f;
                    t = c.sf * t;
                    S = t + 1.0f;
                }
            }
            const float diff = ex_q - S;
            float v = diff * diff;
            if (c.hasW) v = dw_q * v;
            scratch[64 * half + lane] = v;
        }
    } else {
        // slot s in round s, the matrix rows of slot s + kPcAhead loading
        for (int32_t s0 = 0; s0 < nslot; s0 += kPcAhead + 1) {
#pragma unroll
            for (int i = 0; i < kPcAhead + 1; i++) {
                const int32_t s = s0 + i;
                if (s < nslot) {
                    load_slot(mreg[(i + kPcAhead) % (kPcAhead + 1)], s + kPcAhead);
                    float4 *dst = ring + (s & 1) * 2 * kPcGroups * 64;
#pragma unroll
                    for (int j = 0; j < kPcPerProd; j++) {
                        const int32_t rg = pw + kPcProd * j;
                        if (rg < kPcGroups) {
                            const int32_t r = min(s * kPcRows + 4 * rg, Hp - 4);
                            const float4 gv = *reinterpret_cast<const float4 *>(Gs + r);
#pragma unroll
                            for (int hh = 0; hh < 2; hh++) {
                                const float4 mv = mreg[i][2 * j + hh];
                                float4 pr;
                                pr.x = gv.x * mv.x;
                                pr.y = gv.y * mv.y;
                                pr.z = gv.z * mv.z;
                                pr.w = gv.w * mv.w;
                                dst[(hh * kPcGroups + rg) * 64 + lane] = pr;
                            }
                        }
                    }
                    __syncthreads();  // hands slot s to the chain waves
                }
            }
        }
    }
    // the leaf's sum in NumPy's order (the host combines the leaves)
    __syncthreads();
    if (tid == 0) publish(V.out_res + (int64_t)p * V.resW + c.leafOff + leaf, leaf_sum(scratch, len), V.seq);
    clear_next();
}

// grid as k_eval_a's (units x proposals, or XCD-aligned), blocks of
// kPcThreads; leaf-layout windows only
template <int U, bool XCD>
__global__ void __launch_bounds__(kPcThreads) k_eval_pc(View V, int32_t k) {
    extern __shared__ float sm[];
    int32_t u, p;
    int64_t b, nblk;
    if (XCD) {
        const int32_t nu8 = (V.nunits + 7) & ~7;
        u = blockIdx.x % nu8;
        p = blockIdx.x / nu8;
        b = (int64_t)p * V.nunits + u;
        nblk = (int64_t)V.nunits * k;
        if (u >= V.nunits) return;
    } else {
        u = blockIdx.x;
        p = blockIdx.y;
        b = (int64_t)p * gridDim.x + u;
        nblk = (int64_t)gridDim.x * gridDim.y;
    }
    eval_pc_body<U>(V, p, u, sm, b, nblk);
}

// ----------------------------------------------------------- batched windows

// Windows of many engines (replicas) in one launch per kernel: the blocks of
// each active replica's window are numbered consecutively, and a per-kernel
// map gives each block its window (index into the step's windows) and its
// block index within that window.  A block runs the single-engine kernel's
// body on that replica's view, window and proposals.
//
// The map entry, the window record, the replica's view and its proposals are
// copied by the host into device memory (the step block is rewritten at the
// same address every other step, the views once per run).  Every block
// stages what it reads of them into LDS with coherent lane-varying loads:
// uniform reads of copied data would otherwise compile to scalar loads
// through the scalar data cache, which has returned stale values here.
template <class T>
__device__ __forceinline__ void stage(T &dst, const T *src) {
    static_assert(sizeof(T) % 4 == 0, "staged records are whole 32-bit words");
    const uint32_t *s = reinterpret_cast<const uint32_t *>(src);
    uint32_t *d = reinterpret_cast<uint32_t *>(&dst);
    for (int i = threadIdx.x; i < (int)(sizeof(T) / 4); i += blockDim.x) d[i] = load_dev(s + i);
    __syncthreads();
}

// map entry, window record and view of this block of a batched kernel
struct BatchLds {
    int2 m;
    BatchWin W;
    View V;
};

__device__ __forceinline__ void stage_block(BatchLds &S, const int2 *map, const BatchArgs &A) {
    stage(S.m, map + blockIdx.x);
    stage(S.W, A.win + S.m.x);
    stage(S.V, A.views + S.W.rep);
}

// The fields of a replica's view that the batched delta loop reads, copied
// to registers once per workgroup: the view itself sits in global memory,
// where every atomic of the loop could (as far as the compiler knows) change
// it, so reading it in the loop would reload it after each atomic.
struct DeltaCtx {
    int32_t n, nel, C;
    Basis B;
    Grid g[kMaxC];
    int32_t H[kMaxC], dnOff[kMaxC], same[kMaxC];
    int32_t *slot[kMaxC];
    const float *xyz;
    const int32_t *mol, *elem;
};

__device__ __forceinline__ void load_ctx(const View &V, DeltaCtx &X) {
    X.n = V.n;
    X.nel = V.nel;
    X.C = V.C;
    X.B = V.B;
    X.xyz = V.xyz;
    X.mol = V.mol;
    X.elem = V.elem;
#pragma unroll
    for (int ci = 0; ci < kMaxC; ci++) {
        const int cc = ci < X.C ? ci : 0;
        X.g[ci] = V.c[cc].g;
        X.H[ci] = V.c[cc].H;
        X.dnOff[ci] = V.c[cc].dnOff;
        X.same[ci] = ci < X.C ? V.c[cc].sameGrid : 1;
        X.slot[ci] = V.c[cc].slot;
    }
}

// delta_pair and commit_pair on the register copy
__device__ __forceinline__ void delta_pair_x(const DeltaCtx &X, int32_t *out, float so, float sn, int32_t ea,
                                             int32_t eb) {
#pragma unroll
    for (int ci = 0; ci < kMaxC; ci++) {
        if (X.same[ci]) continue;  // also every ci >= C
        const int32_t bo = in_range(so, X.g[ci]) ? bin_of(so, X.g[ci]) : -1;
        const int32_t bn = in_range(sn, X.g[ci]) ? bin_of(sn, X.g[ci]) : -1;
        if (bo == bn) continue;  // -1 and +1 in the same bin (or both out of range) cancel
        if (bo >= 0) {
            const int32_t t = type_index(bo, ea, eb, X.nel, X.H[ci]);
            if (t >= 0) atomicSub(&out[X.dnOff[ci] + t], 1);
        }
        if (bn >= 0) {
            const int32_t t = type_index(bn, ea, eb, X.nel, X.H[ci]);
            if (t >= 0) atomicAdd(&out[X.dnOff[ci] + t], 1);
        }
    }
}

__device__ __forceinline__ void commit_pair_x(const DeltaCtx &X, float s, int32_t ea, int32_t eb, bool intra,
                                              int32_t sign) {
    int32_t b = -1;
#pragma unroll
    for (int ci = 0; ci < kMaxC; ci++) {
        if (ci >= X.C) break;
        if (!X.same[ci]) b = in_range(s, X.g[ci]) ? bin_of(s, X.g[ci]) : -1;
        if (b < 0) continue;
        const int32_t size = X.nel * X.nel * X.H[ci];
        const int32_t f = flat_slot(b, ea, eb, X.nel, X.H[ci]);
        if (f < 0) {
            if (sign > 0) atomicAdd(&X.slot[ci][2 * size], 1);  // past the array: counted, not stored
            continue;
        }
        atomicAdd(&X.slot[ci][(intra ? 0 : size) + f], sign);
    }
}

// One workgroup per (group pb of up to kBP proposals, chunk ch of kChunk
// other atoms) of a replica's window: each other atom's coordinates are read
// once for the group's proposals, whose pair-type deltas are counted in LDS
// (LDSB, when kBP delta blocks fit) and then written out -- with plain stores
// when one chunk covers all atoms (the workgroup owns the blocks), else added
// bin by bin -- or, without LDSB, added straight to the (zeroed) blocks.  The
// replica's pending commit is spread over its workgroups as in k_delta:
// pair-type sums (grid-stride), slot counts (the pb == 0 workgroups, each for
// its chunk), the moved atom's coordinates (one thread).
template <bool PBC, bool ORTHO, bool LDSB>
__device__ __forceinline__ void delta_chunk(const View &V, const BatchWin &W, const Prop *props, int32_t pb,
                                            int32_t ch, int32_t nch, int32_t *bins) {
    // props: this workgroup's kBP proposals (staged in LDS), props[q] for q < np
    const int tid = threadIdx.x;
    DeltaCtx X;
    load_ctx(V, X);
    const int32_t dnStride = V.dnStride, T = V.T, k = W.k;
    const int32_t p0 = pb * kBP, np = min(kBP, k - p0);
    int32_t *out = W.dn + (int64_t)p0 * dnStride;
    const Pending P = pending_of(V, W.P);  // (device window loop: the replica's loop state)
    int32_t qa[kBP], qe[kBP];
    float ox[kBP], oy[kBP], oz[kBP], nx[kBP], ny[kBP], nz[kBP];
#pragma unroll
    for (int q = 0; q < kBP; q++) {
        const Prop &pr = props[q < np ? q : 0];
        qa[q] = q < np ? pr.atom : -1;
        qe[q] = pr.el;
        ox[q] = pr.old[0];
        oy[q] = pr.old[1];
        oz[q] = pr.old[2];
        nx[q] = pr.nw[0];
        ny[q] = pr.nw[1];
        nz[q] = pr.nw[2];
    }
    if (LDSB) {
        for (int32_t s = tid; s < np * dnStride; s += blockDim.x) bins[s] = 0;
        __syncthreads();
    }
    int32_t *acc = LDSB ? bins : out;
    const bool pv = P.valid != 0;
    const int32_t nblk = ((k + kBP - 1) / kBP) * nch;
    if (pv) {
        const int64_t gt = ((int64_t)pb * nch + ch) * blockDim.x + tid;
        const int64_t gs = (int64_t)nblk * blockDim.x;
        for (int32_t ci = 0; ci < X.C; ci++) {
            int32_t *base = V.c[ci].base;
            const int32_t len = T * V.c[ci].H, off = V.c[ci].dnOff;
            for (int64_t s = gt; s < len; s += gs) base[s] += P.dn[off + s];
        }
        if (pb == 0 && ch == 0 && tid == 0)
            for (int d = 0; d < 3; d++) V.xyz[3 * (int64_t)P.atom + d] = P.nw[d];
    }
    const int32_t pmol = pv ? X.mol[P.atom] : 0;
    const int32_t j1 = min(X.n, (ch + 1) * kChunk);
#pragma unroll 2
    for (int32_t j = ch * kChunk + tid; j < j1; j += blockDim.x) {
        float cx = X.xyz[3 * (int64_t)j], cy = X.xyz[3 * (int64_t)j + 1], cz = X.xyz[3 * (int64_t)j + 2];
        const int32_t ej = X.elem[j];
        if (pv) {
            if (pb == 0 && j != P.atom) {
                const float so = pair_s<PBC, ORTHO>(P.old[0], P.old[1], P.old[2], cx, cy, cz, X.B);
                const float sn = pair_s<PBC, ORTHO>(P.nw[0], P.nw[1], P.nw[2], cx, cy, cz, X.B);
                const bool intra = pmol == X.mol[j];
                commit_pair_x(X, so, P.el, ej, intra, -1);
                commit_pair_x(X, sn, P.el, ej, intra, +1);
            }
            if (j == P.atom) {
                cx = P.nw[0];
                cy = P.nw[1];
                cz = P.nw[2];
            }
        }
#pragma unroll
        for (int q = 0; q < kBP; q++) {
            if (q >= np || j == qa[q]) continue;
            // M counts (moved atom a, other j) with diff = c[a] - c[j], before and after
            const float so = pair_s<PBC, ORTHO>(ox[q], oy[q], oz[q], cx, cy, cz, X.B);
            const float sn = pair_s<PBC, ORTHO>(nx[q], ny[q], nz[q], cx, cy, cz, X.B);
            delta_pair_x(X, acc + q * dnStride, so, sn, qe[q], ej);
        }
    }
    if (LDSB) {
        __syncthreads();
        if (nch == 1) {
            for (int32_t s = tid; s < np * dnStride; s += blockDim.x) out[s] = bins[s];
        } else {
            for (int32_t s = tid; s < np * dnStride; s += blockDim.x)
                if (bins[s]) atomicAdd(&out[s], bins[s]);
        }
    }
}

template <bool PBC, bool ORTHO>
__device__ __forceinline__ void delta_chunk_any(const View &V, const BatchWin &W, const Prop *props, int32_t pb,
                                                int32_t ch, int32_t nch, int32_t *bins) {
    if (kBP * V.dnStride <= kBatchLdsInts) delta_chunk<PBC, ORTHO, true>(V, W, props, pb, ch, nch, bins);
    else delta_chunk<PBC, ORTHO, false>(V, W, props, pb, ch, nch, bins);
}

// at most 128 VGPRs (four waves per SIMD): the unconstrained 130 allowed
// three, and 256 replicas ran 8 % faster with four (3.57M against 3.30M
// moves/s; five waves need a register spill).  With two proposals per
// workgroup the bound spills 57 VGPRs to scratch and still wins: 5.41M
// against 4.99M moves/s with three waves and no spill.
__global__ void __launch_bounds__(256, 4) k_delta_batch(BatchArgs A) {
    extern __shared__ int32_t bins[];  // kBP delta blocks (the largest that fits kBatchLdsInts)
    __shared__ BatchLds S;
    __shared__ Prop sp[kBP];
    stage_block(S, A.map_d, A);
    const BatchWin &W = S.W;
    const View &V = S.V;
    const int32_t nch = (V.n + kChunk - 1) / kChunk;
    const int32_t pb = S.m.y / nch, ch = S.m.y % nch;
    {
        // the workgroup's proposals (those past the window's end are not read)
        const int32_t p0 = pb * kBP, np = min(kBP, W.k - p0);
        const uint32_t *src = reinterpret_cast<const uint32_t *>(A.props + W.prop + p0);
        uint32_t *dst = reinterpret_cast<uint32_t *>(sp);
        for (int i = threadIdx.x; i < np * (int)(sizeof(Prop) / 4); i += blockDim.x) dst[i] = load_dev(src + i);
        __syncthreads();
    }
    const Prop *props = sp;
    if (V.isPBC) {
        if (V.ortho) delta_chunk_any<true, true>(V, W, props, pb, ch, nch, bins);
        else delta_chunk_any<true, false>(V, W, props, pb, ch, nch, bins);
    } else {
        delta_chunk_any<false, false>(V, W, props, pb, ch, nch, bins);
    }
}

__global__ void __launch_bounds__(kEvalThreads) k_eval_a_batch(BatchArgs A) {
    extern __shared__ float sm[];
    __shared__ BatchLds S;
    stage_block(S, A.map_a, A);
    const BatchWin &W = S.W;
    const View &V = S.V;
    const int2 m = S.m;
    if (V.C == 1 && V.c[0].rspace) {
        const Fuse F{W.dn_zero, W.zero_len, W.seq, W.k};
        eval_a_body<4, false>(V, W.dn, m.y / V.nunits, m.y % V.nunits, sm, &F, 0, 1);
    } else {
        eval_a_body<4, false>(V, W.dn, m.y / V.nunits, m.y % V.nunits, sm, nullptr, 0, 1);
    }
}

__global__ void __launch_bounds__(256) k_eval_b_batch(BatchArgs A) {
    extern __shared__ float sm[];
    __shared__ BatchLds S;
    stage_block(S, A.map_b, A);
    eval_b_body(S.V, S.W.dn_zero, S.W.zero_len, S.W.seq, S.m.y, S.W.k, sm);
}

// ---------------------------------------------------------------- k_commit


// grid ceil(n/256): picks the proposal to commit (every block makes the same
// decision from the window's totals), then slot counts of the accepted group
// vs every other atom (one atom per thread), pair-type sums += the proposal's
// deltas, and the group's new coordinates (block 0; no block reads the
// group's coordinates).  The group comes from the kernel arguments (single-atom
// windows: old coordinates from the host mirror) or the staging buffers.
template <bool PBC, bool ORTHO>
__global__ void __launch_bounds__(256) k_commit(View V, Inline I, CommitCtl C) {
    __shared__ GroupLds G;
    const int tid = threadIdx.x;
    int32_t p = C.p;
    if (p < 0) {
        // lane q tests proposal q (k <= 64): reject iff newTSE > TSE and
        // random() > tolerance (Engine.py:3308-3317); the first accepted wins
        const int lane = tid % 64;
        const int32_t q = min(lane, C.k - 1);
        const float t = load_dev(V.tse_dev + q);
        const bool acc = lane < C.k && (!(t > C.cur) || !(C.u[q] > C.tol));
        const unsigned long long m = __ballot(acc);
        p = m ? __ffsll((long long)m) - 1 : -1;
    }
    if (blockIdx.x == 0 && tid == 0) {
        // (seq << 32 | p) in one of two slots by window parity: the host checks
        // this window's decision while the next window's commit may have run
        const unsigned long long word = ((unsigned long long)V.seq << 32) | (uint32_t)p;
        __hip_atomic_store(V.out_commit + (V.seq & 1), word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    if (p < 0) return;
    if (I.k > 0) {
        if (tid == 0) {
            const int32_t id = I.atom[p];
            G.g = 1;
            G.id[0] = id;
            G.el[0] = I.el[p];
            G.mol[0] = V.mol[id];
            for (int d = 0; d < 3; d++) {
                G.o[0][d] = I.old[p][d];
                G.w[0][d] = I.nw[p][d];
            }
        }
    } else {
        load_group(V, p, G);
    }
    __syncthreads();
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + tid, gs = (int64_t)gridDim.x * blockDim.x;
    const int32_t *dn = C.dn + (int64_t)p * V.dnStride;
    for (int32_t ci = 0; ci < V.C; ci++) {
        const DevC &c = V.c[ci];
        for (int64_t s = gt; s < (int64_t)V.T * c.H; s += gs) c.base[s] += dn[c.dnOff + s];
    }
    const int32_t g = G.g;
    const int32_t j = (int32_t)gt;
    if (j < V.n) {
        bool member = false;
        for (int32_t a = 0; a < g; a++) member |= (G.id[a] == j);
        if (!member) {
            const float cx = V.xyz[3 * (int64_t)j], cy = V.xyz[3 * (int64_t)j + 1], cz = V.xyz[3 * (int64_t)j + 2];
            const int32_t ej = V.elem[j], mj = V.mol[j];
            for (int32_t a = 0; a < g; a++) {
                const float so = pair_s<PBC, ORTHO>(G.o[a][0], G.o[a][1], G.o[a][2], cx, cy, cz, V.B);
                const float sn = pair_s<PBC, ORTHO>(G.w[a][0], G.w[a][1], G.w[a][2], cx, cy, cz, V.B);
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
            const float so = pair_s<PBC, ORTHO>(G.o[a][0], G.o[a][1], G.o[a][2], G.o[b][0], G.o[b][1], G.o[b][2], V.B);
            const float sn = pair_s<PBC, ORTHO>(G.w[a][0], G.w[a][1], G.w[a][2], G.w[b][0], G.w[b][1], G.w[b][2], V.B);
            const bool intra = G.mol[a] == G.mol[b];
            commit_pair<PBC, ORTHO>(V, so, G.el[a], G.el[b], intra, -1);
            commit_pair<PBC, ORTHO>(V, sn, G.el[a], G.el[b], intra, +1);
        }
        for (int32_t a = tid; a < g; a += blockDim.x)
            for (int d = 0; d < 3; d++) V.xyz[3 * (int64_t)G.id[a] + d] = G.w[a][d];
        if (V.sx && tid == 0) {
            // the sorted copy (kept for single-atom groups): the new position,
            // and its chunk's box grown to hold it
            for (int32_t a = 0; a < g; a++) {
                const int32_t pos = V.spos[G.id[a]];
                V.sx[pos] = make_float4(G.w[a][0], G.w[a][1], G.w[a][2], __int_as_float(G.id[a] * 64 + G.el[a]));
                float *bx = V.cbox + 8 * (int64_t)(pos / kChunkAtoms);
                for (int d = 0; d < 3; d++) {
                    const float f = wrap01(G.w[a][d]);
                    bx[d] = fminf(load_dev(bx + d), f);
                    bx[3 + d] = fmaxf(load_dev(bx + 3 + d), f);
                }
            }
        }
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

int launch_delta(const View &V, const Inline &I, const Pending &P, int32_t k, hipStream_t st) {
    if (V.prune) {
        // single-atom proposals in a periodic box: the chunk walk
        // proposals per block: enough rows of chunks for about kDeltaBlocks
        // blocks (a small box has few chunks, each near most proposals)
        const int64_t want = ((int64_t)k * V.nchunk + kDeltaBlocks - 1) / kDeltaBlocks;
        const int32_t pp = (int32_t)std::max<int64_t>(1, std::min<int64_t>(kPruneP, want));
        const dim3 grid((unsigned)V.nchunk, (unsigned)((k + pp - 1) / pp));
        const int src = V.props ? 2 : I.k > 0 ? 0 : 1;
        if (V.ortho) {
            if (src == 0) k_delta_chunks<true, 0><<<grid, 256, 0, st>>>(V, I, P, k, pp);
            else if (src == 1) k_delta_chunks<true, 1><<<grid, 256, 0, st>>>(V, I, P, k, pp);
            else k_delta_chunks<true, 2><<<grid, 256, 0, st>>>(V, I, P, k, pp);
        } else {
            if (src == 0) k_delta_chunks<false, 0><<<grid, 256, 0, st>>>(V, I, P, k, pp);
            else if (src == 1) k_delta_chunks<false, 1><<<grid, 256, 0, st>>>(V, I, P, k, pp);
            else k_delta_chunks<false, 2><<<grid, 256, 0, st>>>(V, I, P, k, pp);
        }
        FRMC_CHECK_LAUNCH();
        return FRMC_OK;
    }
    const dim3 grid(blocks_for(V.n, 256), (unsigned)((k + kDP - 1) / kDP));
    if (V.props) {
        if (!V.isPBC) k_delta<false, false, true><<<grid, 256, 0, st>>>(V, I, P, k);
        else if (V.ortho) k_delta<true, true, true><<<grid, 256, 0, st>>>(V, I, P, k);
        else k_delta<true, false, true><<<grid, 256, 0, st>>>(V, I, P, k);
    } else if (V.isPBC) {
        if (V.ortho) k_delta<true, true, false><<<grid, 256, 0, st>>>(V, I, P, k);
        else k_delta<true, false, false><<<grid, 256, 0, st>>>(V, I, P, k);
    } else {
        k_delta<false, false, false><<<grid, 256, 0, st>>>(V, I, P, k);
    }
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

size_t eval_lds_bytes(const View &V) {
    // the leaf's 128 squared deviations reuse the pairwise scratch
    return sizeof(float) * ((size_t)((V.Hmax + 15) & ~15) + std::max<size_t>(V.maxLeaves + kStack, 128) +
                            (size_t)2 * V.maxLeaves);
}

template <int U, bool XCD, bool IL>
static void launch_k_eval_a(const View &V, int32_t k, hipStream_t st) {
    const size_t shm = eval_lds_bytes(V);
    if (shm > 48 * 1024)
        (void)hipFuncSetAttribute((const void *)k_eval_a<U, XCD, IL>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)shm);
    if (XCD)
        k_eval_a<U, XCD, IL><<<dim3((unsigned)((V.nunits + 7) & ~7) * (unsigned)k), kEvalThreads, shm, st>>>(V, k);
    else
        k_eval_a<U, XCD, IL><<<dim3((unsigned)V.nunits, (unsigned)k), kEvalThreads, shm, st>>>(V, k);
}

template <bool IL>
static void launch_k_eval_a_il(const View &V, int32_t k, hipStream_t st, bool u8, bool xcd) {
    if (xcd) {
        if (u8) launch_k_eval_a<8, true, IL>(V, k, st); else launch_k_eval_a<4, true, IL>(V, k, st);
    } else {
        if (u8) launch_k_eval_a<8, false, IL>(V, k, st); else launch_k_eval_a<4, false, IL>(V, k, st);
    }
}

template <int U, bool XCD>
static void launch_k_eval_pc(const View &V, int32_t k, hipStream_t st) {
    const size_t shm = pc_lds_bytes(V);
    (void)hipFuncSetAttribute((const void *)k_eval_pc<U, XCD>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (XCD)
        k_eval_pc<U, XCD><<<dim3((unsigned)((V.nunits + 7) & ~7) * (unsigned)k), kPcThreads, shm, st>>>(V, k);
    else
        k_eval_pc<U, XCD><<<dim3((unsigned)V.nunits, (unsigned)k), kPcThreads, shm, st>>>(V, k);
}

// shape s of k_eval_a: bins per thread per G(r) pass U = 8 if s & 4 (and the
// grid has more than 1,024 bins) else 4; s & 8: XCD-aligned units; s & 32:
// the interleaved S(Q) matrix (s & 16, the unit layout, is the caller's);
// s & 64: the split chain of k_eval_pc (leaf layout, interleaved matrix;
// U = 2 or 4 bins per pass for its 1,024 threads).  Every shape gives the
// same results.
int launch_eval(const View &V, int32_t k, hipStream_t st, int32_t shape) {
    if (shape & 64) {
        const bool u4 = (shape & 4) && V.Hmax > 2 * kPcThreads;
        if (shape & 8) {
            if (u4) launch_k_eval_pc<4, true>(V, k, st); else launch_k_eval_pc<2, true>(V, k, st);
        } else {
            if (u4) launch_k_eval_pc<4, false>(V, k, st); else launch_k_eval_pc<2, false>(V, k, st);
        }
        FRMC_CHECK_LAUNCH();
        return FRMC_OK;
    }
    const bool u8 = (shape & 4) && V.Hmax > 4 * kEvalThreads;
    if (shape & 32) launch_k_eval_a_il<true>(V, k, st, u8, (shape & 8) != 0);
    else launch_k_eval_a_il<false>(V, k, st, u8, (shape & 8) != 0);
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

int launch_eval_b(const View &V, int32_t k, hipStream_t st) {
    const size_t shm_b = eval_b_lds_bytes(V);
    if (shm_b > 48 * 1024)
        (void)hipFuncSetAttribute((const void *)k_eval_b, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm_b);
    k_eval_b<<<(unsigned)k, 256, shm_b, st>>>(V, k);
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

int launch_batch(const BatchArgs &A, int32_t nd, int32_t na, int32_t nb, size_t shm_d, size_t shm_a, size_t shm_b,
                 hipStream_t st) {
    k_delta_batch<<<(unsigned)nd, 256, shm_d, st>>>(A);
    FRMC_CHECK_LAUNCH();
    if (shm_a > 48 * 1024)
        (void)hipFuncSetAttribute((const void *)k_eval_a_batch, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)shm_a);
    k_eval_a_batch<<<(unsigned)na, kEvalThreads, shm_a, st>>>(A);
    FRMC_CHECK_LAUNCH();
    if (nb == 0) return FRMC_OK;  // every window's total came from k_eval_a_batch
    if (shm_b > 48 * 1024)
        (void)hipFuncSetAttribute((const void *)k_eval_b_batch, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)shm_b);
    k_eval_b_batch<<<(unsigned)nb, 256, shm_b, st>>>(A);
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

size_t eval_b_lds_bytes(const View &V) {
    return sizeof(float) * (V.Mmax + V.maxLeaves + kStack + (size_t)2 * V.maxLeaves);
}

int launch_commit(const View &V, const Inline &I, const CommitCtl &C, hipStream_t st) {
    const unsigned g = blocks_for(V.n, 256);
    if (V.isPBC) {
        if (V.ortho) k_commit<true, true><<<g, 256, 0, st>>>(V, I, C);
        else k_commit<true, false><<<g, 256, 0, st>>>(V, I, C);
    } else {
        k_commit<false, false><<<g, 256, 0, st>>>(V, I, C);
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
