// window_loop.hip -- k_loop_step: one step of the device window loop for
// single-atom translation runs, one workgroup (k_loop_step_batch: one per
// replica of a batched run).
//
// The host queues, per window, k_loop_step, k_delta and the evaluation on the
// engine stream, for a batch of windows of one size k, and synchronises once
// per batch.  k_loop_step
//   1. walks the totals of the window in flight in proposal order with the
//      reference's accept rule (Engine.py:3308-3317): a proposal is accepted
//      when its total is not worse than the current one, or when it is worse
//      and its acceptance draw, rounded to float32, is not above the
//      tolerance; proposals after the first accepted one are discarded.  It
//      logs the walked proposals, records the accepted move (new state, the
//      commit the next k_delta applies, the real coordinates), and
//   2. advances both MT19937 streams from the window's start to just after
//      the last walked proposal's draws (the acceptance draw of an accepted
//      proposal counts only if its total was worse), the state the
//      sequential loop holds there, and
//   3. draws the next window's k proposals (fewer when fewer moves are
//      left) from there exactly as the host loop does
//      (RandomSelector.select_index, Selectors/RandomSelectors.py:87-94;
//      generate_random_vector, Core/Collection.py:674-701; one acceptance
//      draw each), and writes them into the proposal table k_delta reads.
// Per stream the workgroup holds up to four consecutive state blocks in LDS
// (block b+1 twisted out of block b, 3 dependent phases of independent
// words).  The Python stream's word offsets depend on randrange's rejections:
// wave 0 marks the usable words of 64-word tiles with one ballot each and
// finds the proposals in a tile with scalar bit operations; then each
// proposal is made by its own lane from its words.
#include <algorithm>

#include "draws.h"
#include "engine.h"

namespace frmc {

constexpr int kLoopThreads = 256;
constexpr int kLoopRing = 4;  // state blocks held per stream

// block n = twist(block o) (the sequential recurrence: word i reads words i
// and i+1 of o and word i+397 of o for i < 227, word i-227 of n after)
__device__ void twist_block(const uint32_t *o, uint32_t *n) {
    const int t = threadIdx.x;
    for (int i = t; i < 227; i += kLoopThreads) n[i] = o[i + 397] ^ mt_mix(o[i], o[i + 1]);
    __syncthreads();
    for (int i = 227 + t; i < 454; i += kLoopThreads) n[i] = n[i - 227] ^ mt_mix(o[i], o[i + 1]);
    __syncthreads();
    for (int i = 454 + t; i < kMtN; i += kLoopThreads) n[i] = n[i - 227] ^ mt_mix(o[i], i == kMtN - 1 ? n[0] : o[i + 1]);
    __syncthreads();
}

__device__ __forceinline__ float res_value(const unsigned long long *w) {
    return __uint_as_float((uint32_t)*w);
}

// Everything the step reads from memory is loaded in one parallel phase at
// its start (state blocks, the walked window's results and proposals, the
// combine trees) and once more for the drawn atoms (their coordinates), and
// everything it writes is stored at its end: the step sits between two
// windows on the stream, so its latency is what it costs.
__device__ void loop_step_body(const LoopArgs &A) {
    __shared__ uint32_t ring[2][kLoopRing][kMtN];
    extern __shared__ float s_res[];  // [wk][rs]: result words, then the combine tree's nodes
    __shared__ int4 s_node[kLoopNodes];
    __shared__ Prop s_prop[kLoopMaxK];
    __shared__ float s_nreal[kLoopMaxK][3];
    __shared__ int32_t s_gid[kLoopMaxK], s_slot[kLoopMaxK], s_pre[kLoopMaxK], s_post[kLoopMaxK];
    __shared__ float s_u[kLoopMaxK], s_tse[kLoopMaxK];
    __shared__ float s_chi[kLoopMaxK][kMaxC];
    __shared__ int32_t s_r[kLoopMaxK];
    __shared__ int32_t s_first, s_scan_p, s_scan_j;
    // the accepted move not yet committed on the device (its new box and
    // real coordinates replace the atom's stored ones in the draws)
    __shared__ int32_t s_pv, s_patom, s_preal_set;
    __shared__ float s_pbox[3], s_preal[3];

    __shared__ LoopHead s_head;
    __shared__ int32_t s_pos[2];

    LoopState *L = A.L;
    const int tid = threadIdx.x;
    // ------------------------------------------------ 0. loads
    // Every load of this phase is issued before the first wait: lane-varying
    // addresses (vector loads, coherent with what earlier kernels wrote),
    // into registers, then into LDS.
    {
        constexpr int kHeadWords = (int)(sizeof(LoopHead) / 4);
        const uint32_t *pyw = (const uint32_t *)&L->py, *npw = (const uint32_t *)&L->np;  // 624 words + position
        // indexes clamped into the arrays so that every lane loads (no
        // per-lane branches around the loads, whose merges would wait)
        uint32_t vp[3], vn[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int i = min(tid + kLoopThreads * j, kMtN);
            vp[j] = pyw[i];
            vn[j] = npw[i];
        }
        const uint32_t vh = ((const uint32_t *)&L->h)[min(tid, kHeadWords - 1)];
        // the walked window: result values (low words), proposals and draws
        const int32_t nres = A.wk * A.resW;
        const int32_t rlast = max(nres - 1, 0);
        const uint32_t *rw = (const uint32_t *)A.res;
        constexpr int kRes = 8;
        uint32_t vr[kRes];
#pragma unroll
        for (int j = 0; j < kRes; j++) vr[j] = rw[2 * min(tid + kLoopThreads * j, rlast)];
        int4 vg = make_int4(0, 0, 0, 0);
        if (A.nnodes > 0) vg = A.nodes[min(tid, A.nnodes - 1)];
        const int32_t pw = min(tid, kLoopMaxK - 1);
        const Prop pr = A.props[pw];
        const int32_t gid = L->gid[pw], slot = L->slot[pw], pre = L->py_pre[pw], post = L->py_post[pw];
        const float u = L->uacc[pw];
        float nr[3];
        for (int d = 0; d < 3; d++) nr[d] = L->nreal[pw][d];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int i = tid + kLoopThreads * j;
            if (i < kMtN) {
                ring[0][0][i] = vp[j];
                ring[1][0][i] = vn[j];
            } else if (i == kMtN) {
                s_pos[0] = (int32_t)vp[j];
                s_pos[1] = (int32_t)vn[j];
            }
        }
        if (tid < kHeadWords) ((uint32_t *)&s_head)[tid] = vh;
        // result word w of proposal p goes to row p, word w
#pragma unroll
        for (int j = 0; j < kRes; j++) {
            const int32_t i = tid + kLoopThreads * j;
            if (i < nres) s_res[(i / A.resW) * A.rs + i % A.resW] = __uint_as_float(vr[j]);
        }
        for (int32_t i = tid + kLoopThreads * kRes; i < nres; i += kLoopThreads)
            s_res[(i / A.resW) * A.rs + i % A.resW] = __uint_as_float(rw[2 * i]);
        if (tid < A.nnodes) s_node[tid] = vg;
        if (tid < A.wk) {
            s_prop[tid] = pr;
            s_gid[tid] = gid;
            s_slot[tid] = slot;
            s_pre[tid] = pre;
            s_post[tid] = post;
            s_u[tid] = u;
            for (int d = 0; d < 3; d++) s_nreal[tid][d] = nr[d];
        }
    }
    if (tid == 0) s_preal_set = 0;
    __syncthreads();
    // proposals of the window in flight (fewer than its size when the run's
    // last moves did not fill it)
    const int32_t kw = A.wk > 0 ? min(s_head.kw, A.wk) : 0;
    if (tid == 0) {
        s_first = kw;
        s_pv = s_head.pend.valid;
        s_patom = s_head.pend.atom;
        for (int d = 0; d < 3; d++) s_pbox[d] = s_head.pend.nw[d];
    }
    // per stream: position in the start block, start block, blocks held [., hi)
    int32_t pos[2] = {s_pos[0], s_pos[1]};
    int32_t sb[2] = {0, 0}, hi[2] = {1, 1};
    int64_t step = A.reset ? 0 : s_head.step;
    int32_t nacc = A.reset ? 0 : s_head.nacc;
    int32_t ntol = A.reset ? 0 : s_head.ntol;
    const float cur = s_head.cur_tse;
    __syncthreads();

    // ------------------------------------------------ 1. walk
    // (its results are stored at the end, with the draws: a barrier waits
    // for every global store issued before it)
    int32_t used[2] = {0, 0};  // words of each stream the walked proposals used
    const int64_t step0 = step;
    int32_t a = -1, walked = 0;  // first accepted proposal, proposals walked
    if (A.wk > 0) {
        if (A.leafmode) {
            // the leaf combines, one tree level at a time over all proposals
            for (int32_t lv = 1; lv <= A.nlev; lv++) {
                const int32_t n0 = A.lev_off[lv - 1], nn = A.lev_off[lv] - n0;
                for (int32_t t = tid; t < kw * nn; t += kLoopThreads) {
                    const int32_t p = t / nn;
                    const int4 nd = s_node[n0 + t % nn];
                    float *row = s_res + p * A.rs;
                    row[nd.z] = row[nd.x] + row[nd.y];
                }
                __syncthreads();
            }
        }
        if (tid < kw) {
            const int32_t p = tid;
            const float *rv = s_res + p * A.rs;
            float tse = 0.0f;
            if (!A.leafmode) {
                tse = rv[0];
                for (int ci = 0; ci < A.C; ci++) s_chi[p][ci] = rv[1 + ci];
            } else {
                // (Engine.py:3026-3029: the total is np.sum of the chi-squares from 0)
                for (int ci = 0; ci < A.C; ci++) {
                    const float chi = rv[A.root[ci]];
                    s_chi[p][ci] = chi;
                    tse = tse + chi;
                }
            }
            s_tse[p] = tse;
            const bool worse = tse > cur;
            if (!worse || !(s_u[p] > A.tol)) atomicMin(&s_first, p);
        }
        __syncthreads();
        a = s_first < kw ? s_first : -1;
        walked = a >= 0 ? a + 1 : kw;
        const bool worse_a = a >= 0 && s_tse[a] > cur;
        if (a >= 0) {
            used[0] = worse_a ? s_post[a] : s_pre[a];
            used[1] = 6 * (a + 1);
        } else {
            used[0] = kw > 0 ? s_post[kw - 1] : 0;
            used[1] = 6 * kw;
        }
        // the accepted move replaces its atom's stored coordinates in the draws
        if (tid == 0) {
            s_pv = a >= 0 ? 1 : 0;
            if (a >= 0) {
                s_patom = s_prop[a].atom;
                s_preal_set = 1;
                for (int d = 0; d < 3; d++) {
                    s_pbox[d] = s_prop[a].nw[d];
                    s_preal[d] = s_nreal[a][d];
                }
            }
        }
        if (a >= 0) {
            nacc++;
            if (worse_a) ntol++;
        }
        step += walked;
        __syncthreads();
    }

    // ------------------------------------------------ 2. both streams to the next window's start
    for (int s = 0; s < 2; s++) {
        if (used[s] > 0) {
            const int32_t at = pos[s] + used[s];
            sb[s] = (at - 1) / kMtN;
            pos[s] = at - sb[s] * kMtN;  // in [1, 624]: the host twists only when it draws past a block
        }
        while (hi[s] <= sb[s]) {
            twist_block(ring[s][(hi[s] - 1) & 3], ring[s][hi[s] & 3]);
            hi[s]++;
        }
    }

    // ------------------------------------------------ 3. draw the next window
    // word o of a stream counted from the window's start
    auto word = [&](int s, int32_t o) -> uint32_t {
        const int32_t r = pos[s] + o;
        return mt_temper(ring[s][(sb[s] + r / kMtN) & 3][r % kMtN]);
    };
    bool words_err = false;  // the draws need more Python-stream words than the ring holds
    const int32_t dk = A.dk > 0 ? (int32_t)min((int64_t)A.dk, max((int64_t)0, A.left - step)) : 0;
    if (dk > 0) {
        // NumPy stream: 6 words (three uniforms) per proposal
        const int32_t last_np = sb[1] + (pos[1] + 6 * dk - 1) / kMtN;
        while (hi[1] <= last_np) {
            twist_block(ring[1][(hi[1] - 1) & 3], ring[1][hi[1] & 3]);
            hi[1]++;
        }
        // Python stream: per proposal randrange(G) (one word per try, taken
        // when its top nbits bits are below G), then two words for the
        // amplitude and two for the acceptance draw
        if (tid == 0) {
            s_scan_p = 0;
            s_scan_j = 0;
        }
        __syncthreads();
        for (;;) {
            const int32_t avail = (hi[0] - sb[0]) * kMtN - pos[0];
            if (tid < 64) {
                // one ballot marks the usable words of a 64-word tile; the
                // proposals in it are then found with scalar bit operations
                int32_t p = s_scan_p, j = s_scan_j;
                while (p < dk && j + 68 <= avail) {
                    const uint32_t w = word(0, j + tid);
                    const unsigned long long m = __ballot((w >> (32 - A.nbits)) < (uint32_t)A.G);
                    const int32_t base = j;
                    j = base + 64;
                    int32_t off = 0;
                    while (p < dk && off < 64) {
                        const unsigned long long mm = m & (~0ull << off);
                        if (mm == 0) {
                            j = base + 64;  // the rest of the tile: rejected tries
                            break;
                        }
                        const int32_t r = base + __ffsll(mm) - 1;
                        if (tid == 0) s_r[p] = r;
                        p++;
                        j = r + 5;
                        off = j - base;
                    }
                }
                if (tid == 0) {
                    s_scan_p = p;
                    s_scan_j = j;
                }
            }
            __syncthreads();
            if (s_scan_p >= dk) break;
            if (hi[0] - sb[0] >= kLoopRing) {  // (uniform: every thread leaves here)
                words_err = true;
                break;
            }
            twist_block(ring[0][(hi[0] - 1) & 3], ring[0][hi[0] & 3]);
            hi[0]++;
        }
        if (!words_err && tid < dk) {
            const int32_t p = tid, r = s_r[p];
            const int32_t gi = (int32_t)(word(0, r) >> (32 - A.nbits));
            const int32_t atom = A.ident ? gi : A.gatoms[A.goff[gi]];
            // the atom's data first (one memory round trip), then the draws
            float ro[3], bo[3];
            for (int d = 0; d < 3; d++) {
                ro[d] = A.real[3 * (int64_t)atom + d];
                bo[d] = A.xyz[3 * (int64_t)atom + d];
            }
            const int32_t el = A.elem[atom];
            const int32_t slot = A.spos ? A.spos[atom] : -1;
            const double ua = mt_uniform(word(0, r + 1), word(0, r + 2));
            const double uc = mt_uniform(word(0, r + 3), word(0, r + 4));
            double un[3];
            for (int d = 0; d < 3; d++) un[d] = mt_uniform(word(1, 6 * p + 2 * d), word(1, 6 * p + 2 * d + 1));
            float v[3], norm;
            if (!random_direction(un, v, norm)) {
                L->h.error = kLoopErrDirection;
                for (int d = 0; d < 3; d++) v[d] = 0.0f;
            } else {
                scale_direction(v, norm, ua, A.minAmp, A.maxAmp);
            }
            const bool pend = s_pv && atom == s_patom;
            Prop q;
            q.atom = atom;
            q.el = el;
            float rn[3];
            for (int d = 0; d < 3; d++) {
                if (pend) {
                    bo[d] = s_pbox[d];
                    if (s_preal_set) ro[d] = s_preal[d];
                }
                rn[d] = ro[d] + v[d];
                q.old[d] = bo[d];
            }
            to_box(A.recip, rn, q.nw);
            A.props[p] = q;
            L->gid[p] = gi;
            L->slot[p] = slot;
            L->py_pre[p] = r + 3;
            L->py_post[p] = r + 5;
            L->uacc[p] = (float)uc;
            for (int d = 0; d < 3; d++) L->nreal[p][d] = rn[d];
        }
    }

    // ------------------------------------------------ 4. the walk's and the streams' stores
    if (A.wk > 0 && tid < walked && A.glog) {
        A.glog[step0 + tid] = s_gid[tid];
        A.alog[step0 + tid] = tid == a ? 1 : 0;
        A.tlog[step0 + tid] = tid == a ? s_tse[tid] : cur;
    }
    for (int i = tid; i < kMtN; i += kLoopThreads) {
        if (sb[0] > 0) L->py.mt[i] = ring[0][sb[0] & 3][i];
        if (sb[1] > 0) L->np.mt[i] = ring[1][sb[1] & 3][i];
    }
    if (tid == 0) {
        if (A.wk > 0) {
            // the commit the next k_delta applies, the accepted state
            Pending P;
            P.valid = 0;
            P.atom = -1;
            P.el = 0;
            P.slot = -1;
            P.dn = A.win_dn;
            for (int d = 0; d < 3; d++) P.old[d] = P.nw[d] = 0.0f;
            if (a >= 0) {
                const Prop &q = s_prop[a];
                P.valid = 1;
                P.atom = q.atom;
                P.el = q.el;
                P.slot = s_slot[a];
                P.dn = A.win_dn + (int64_t)a * A.dnStride;
                LoopAcc acc;
                acc.atom = q.atom;
                for (int d = 0; d < 3; d++) {
                    P.old[d] = q.old[d];
                    P.nw[d] = q.nw[d];
                    acc.box[d] = q.nw[d];
                    acc.real[d] = s_nreal[a][d];
                    A.real[3 * (int64_t)q.atom + d] = acc.real[d];
                }
                A.acc[nacc - 1] = acc;
                L->h.cur_tse = s_tse[a];
                for (int ci = 0; ci < A.C; ci++) L->h.cur_chi2[ci] = s_chi[a][ci];
            }
            L->h.pend = P;
        }
        L->py.pos = pos[0];
        L->np.pos = pos[1];
        L->h.kw = dk;
        L->h.nwin = (A.reset ? 0 : s_head.nwin) + (dk > 0 ? 1 : 0);
        L->h.neval = (A.reset ? 0 : s_head.neval) + dk;
        L->h.step = step;
        L->h.nacc = nacc;
        L->h.ntol = ntol;
        if (words_err) L->h.error = kLoopErrWords;
    }
}

__global__ void __launch_bounds__(kLoopThreads) k_loop_step(LoopArgs A) { loop_step_body(A); }

// Step s of a batch of M windows for every replica (one workgroup each):
// the replica's arguments are copied to LDS, with this step's window (walk
// of window s-1, draw of window s) filled in.
__global__ void __launch_bounds__(kLoopThreads) k_loop_step_batch(const LoopArgs *args, int32_t s, int32_t M) {
    __shared__ LoopArgs sA;
    constexpr int kWords = (int)(sizeof(LoopArgs) / 4);
    static_assert(sizeof(LoopArgs) % 4 == 0 && kWords <= kLoopThreads, "LoopArgs copy");
    const uint32_t *src = (const uint32_t *)(args + blockIdx.x);
    if (threadIdx.x < kWords) ((uint32_t *)&sA)[threadIdx.x] = src[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        sA.reset = s == 0 ? 1 : 0;
        sA.wk = s > 0 ? sA.kr : 0;
        sA.dk = s < M ? sA.kr : 0;
        sA.leafmode = 0;
        sA.win_dn = sA.dnb[(sA.c0 + s + 1) & 1];  // window s-1's buffer
    }
    __syncthreads();
    loop_step_body(sA);
}

int launch_loop_step_batch(const LoopArgs *args, int32_t R, int32_t s, int32_t M, size_t shm, hipStream_t st) {
    FRMC_REQUIRE(R >= 1 && s >= 0 && s <= M && M <= kLoopWindows, "bad loop batch step");
    FRMC_REQUIRE(shm <= sizeof(float) * kLoopMaxK * kLoopResW, "bad loop batch LDS");
    static const hipError_t attr = hipFuncSetAttribute(
        (const void *)k_loop_step_batch, hipFuncAttributeMaxDynamicSharedMemorySize,
        (int)(sizeof(float) * kLoopMaxK * (kLoopResW + kLoopNodes)));
    FRMC_REQUIRE(attr == hipSuccess, "k_loop_step_batch: dynamic LDS not available");
    k_loop_step_batch<<<(unsigned)R, kLoopThreads, std::max<size_t>(shm, 4), st>>>(args, s, M);
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

int launch_loop_step(const LoopArgs &A, hipStream_t st) {
    FRMC_REQUIRE(A.wk >= 0 && A.wk <= kLoopMaxK && A.dk >= 0 && A.dk <= kLoopMaxK, "loop window out of range");
    FRMC_REQUIRE(A.dk == 0 || (A.G >= 1 && A.nbits == bit_length(A.G)), "bad group count");
    FRMC_REQUIRE(A.nnodes >= 0 && A.nnodes <= kLoopNodes && A.nlev >= 0 && A.nlev <= kLoopLevels &&
                     A.rs >= A.resW + A.nnodes && A.resW <= kLoopResW,
                 "bad combine trees");
    static const hipError_t attr = hipFuncSetAttribute(
        (const void *)k_loop_step, hipFuncAttributeMaxDynamicSharedMemorySize,
        (int)(sizeof(float) * kLoopMaxK * (kLoopResW + kLoopNodes)));
    FRMC_REQUIRE(attr == hipSuccess, "k_loop_step: dynamic LDS not available");
    const size_t shm = sizeof(float) * (size_t)std::max(A.wk, 1) * A.rs;
    k_loop_step<<<1, kLoopThreads, shm, st>>>(A);
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

}  // namespace frmc
