// engine.hip -- C ABI of the device-resident engine: creation from a host
// descriptor, full recompute, evaluation of a window of k proposals, commit,
// state transfer and per-kernel timing.  The kernels are in
// engine_kernels.hip, the native move driver in driver.hip.
//
// A window: the host writes the proposals into one of two pinned staging
// buffers (copied to the device unless the proposals fit in the kernel
// arguments), launches k_delta, k_eval_a and k_eval_b on the engine stream,
// and polls the result words k_eval_b writes into pinned memory.  An accepted single-atom move of the native loop is committed by
// the next window's k_delta (or by flush_commit before anything else touches
// the state); other commits are launched behind the window without waiting.
// The staging buffer a commit reads is not rewritten until the next window
// has completed.
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <time.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "draws.h"
#include "engine.h"

namespace frmc {

View make_view(frmc_engine *e, int s) {
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
    for (int i = 0; i < kMaxC; i++) V.c[i] = e->c[i];
    V.dnStride = e->dnStride;
    V.Hmax = e->Hmax;
    V.Mmax = e->Mmax;
    V.maxLeaves = e->maxLeaves;
    V.nunits = e->nunits;
    V.pgroup = e->stage[s].d_pgroup;
    V.poff = e->stage[s].d_poff;
    V.patoms = e->stage[s].lists ? e->stage[s].d_patoms : nullptr;
    V.pold = e->stage[s].d_pold;
    V.pnew = e->stage[s].d_pnew;
    V.dn = e->dnbuf[e->cur_dn];
    V.dn_zero = e->dnbuf[1 - e->cur_dn];
    V.zero_len = (int64_t)e->dn_used[1 - e->cur_dn] * e->dnStride;
    V.cnt = e->cnt;
    V.sqbuf = e->sqbuf;
    V.chi2d = e->chi2d;
    V.out_res = e->h_res;
    V.resW = e->resW;
    V.leafmode = 0;
    V.leafunits = e->leaf_units ? 1 : 0;
    V.tse_dev = e->tse_dev;
    V.out_commit = e->h_commit;
    V.seq = e->seq;
    const bool sorted = e->cells && !e->cells_stale;
    V.sx = sorted ? e->sx : nullptr;
    V.cbox = e->cbox;
    V.spos = e->spos;
    V.nchunk = e->nchunk;
    V.prune = sorted && !e->stage[s].lists ? 1 : 0;  // the chunk walk takes single-atom groups
    for (int d = 0; d < 3; d++) V.hsp[d] = e->hsp[d];
    V.wcut = e->wcut;
    V.pend_dev = nullptr;
    V.props = nullptr;
    return V;
}


static hipEvent_t take_event(frmc_engine *e) {
    if (!e->event_pool.empty()) {
        hipEvent_t ev = e->event_pool.back();
        e->event_pool.pop_back();
        return ev;
    }
    hipEvent_t ev = nullptr;
    (void)hipEventCreate(&ev);
    return ev;
}

KTimer::KTimer(frmc_engine *e_, int kid_) : e(e_), kid(kid_) {
    if (e->profile && e->seq % (uint32_t)e->profile_every == 0) {
        a = take_event(e);
        (void)hipEventRecord(a, e->st);
    }
}

KTimer::~KTimer() {
    if (e->profile && a) {
        hipEvent_t b = take_event(e);
        (void)hipEventRecord(b, e->st);
        e->pending.push_back({kid, a, b});
    }
}

// fold finished event pairs into the per-kernel totals; unfinished pairs stay
void collect_times(frmc_engine *e) {
    std::vector<frmc_engine::Pending> keep;
    for (auto &p : e->pending) {
        if (hipEventQuery(p.b) != hipSuccess) {
            keep.push_back(p);
            continue;
        }
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            e->kernel_ms[p.kid] += ms;
            e->kernel_launches[p.kid]++;
        }
        e->event_pool.push_back(p.a);
        e->event_pool.push_back(p.b);
    }
    e->pending.swap(keep);
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

// pinned, device-mapped, coherent host memory (device accesses are not cached)
template <typename T>
static hipError_t halloc(frmc_engine *e, T **p, size_t count) {
    hipError_t err = hipHostMalloc((void **)p, sizeof(T) * (count ? count : 1),
                                   hipHostMallocMapped | hipHostMallocCoherent);
    if (err == hipSuccess) {
        memset((void *)*p, 0, sizeof(T) * (count ? count : 1));
        e->host_allocs.push_back((void *)*p);
    }
    return err;
}

// ------------------------------------------------------- sorted atom copy

// frmc_set_delta_cells: -1 chunk walk for periodic single-atom engines of at
// least kCellsMinAtoms atoms, 0 never, 1 for every periodic single-atom engine
static int g_cells_mode = -1;
constexpr int32_t kCellsMinAtoms = 4096;
// box coordinates beyond this magnitude lose the float32 resolution the
// chunk test's margin assumes (the sorted copy is then not used)
constexpr float kCellsMaxCoord = 64.0f;

// Sort the atoms by a grid of cells of about kChunkAtoms atoms each (cell
// index z-major over the wrapped fractional coordinates), so that a chunk of
// kChunkAtoms consecutive atoms spans one or two cells; upload the copy, the
// chunk boxes and each atom's position.  The stream is drained first (the
// kernels of earlier windows read the copy).  A coordinate beyond
// kCellsMaxCoord turns the copy off for good.
static float tag_float(int32_t t) {
    float f;
    memcpy(&f, &t, sizeof f);
    return f;
}

static int build_cells(frmc_engine *e) {
    const int32_t n = e->n;
    for (size_t q = 0; q < e->h_box.size(); q++)
        if (!(fabsf(e->h_box[q]) <= kCellsMaxCoord)) {
            e->cells = false;
            return FRMC_OK;
        }
    FRMC_CHECK_HIP(hipStreamSynchronize(e->st));
    // lattice plane spacings h_i = V / |a_j x a_k| (basis rows a_i)
    double a[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) a[r][c] = e->B.b[3 * r + c];
    auto cross = [](const double *u, const double *v, double *w) {
        w[0] = u[1] * v[2] - u[2] * v[1];
        w[1] = u[2] * v[0] - u[0] * v[2];
        w[2] = u[0] * v[1] - u[1] * v[0];
    };
    double cx[3][3];
    cross(a[1], a[2], cx[0]);
    cross(a[2], a[0], cx[1]);
    cross(a[0], a[1], cx[2]);
    const double vol = fabs(a[0][0] * cx[0][0] + a[0][1] * cx[0][1] + a[0][2] * cx[0][2]);
    const double L = cbrt((double)kChunkAtoms * vol / n);  // edge of a cell of about kChunkAtoms atoms
    int32_t nc[3];
    double hmax = 0.0;
    for (int i = 0; i < 3; i++) {
        const double h = vol / sqrt(cx[i][0] * cx[i][0] + cx[i][1] * cx[i][1] + cx[i][2] * cx[i][2]);
        e->hsp[i] = (float)h;
        hmax = std::max(hmax, h);
        nc[i] = (int32_t)std::min(256.0, std::max(1.0, std::floor(h / L + 0.5)));
    }
    float maxD = 0.0f;
    for (int ci = 0; ci < e->C; ci++) maxD = std::max(maxD, e->c[ci].g.maxD);
    e->wcut = (float)(maxD * (1.0 + 1e-5) + 0.01 + 1e-4 * hmax);
    // counting sort by cell
    const int64_t ncell = (int64_t)nc[0] * nc[1] * nc[2];
    std::vector<int32_t> cell(n), start(ncell + 1, 0);
    std::vector<float> wf(3 * (size_t)n);
    for (int32_t i = 0; i < n; i++) {
        int32_t ci[3];
        for (int d = 0; d < 3; d++) {
            const float x = e->h_box[3 * (size_t)i + d];
            const float f = x - std::floor(x);
            wf[3 * (size_t)i + d] = f;
            ci[d] = std::min((int32_t)(f * nc[d]), nc[d] - 1);
        }
        cell[i] = (ci[2] * nc[1] + ci[1]) * nc[0] + ci[0];
        start[cell[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; c++) start[c + 1] += start[c];
    const int32_t nchunk = (n + kChunkAtoms - 1) / kChunkAtoms;
    std::vector<float4> sx((size_t)nchunk * kChunkAtoms, make_float4(0.f, 0.f, 0.f, tag_float(-1)));
    std::vector<float> cbox((size_t)nchunk * 8, 0.0f);
    e->h_spos.assign(n, 0);
    for (int32_t i = 0; i < n; i++) {
        const int32_t pos = start[cell[i]]++;
        e->h_spos[i] = pos;
        const float *x = &e->h_box[3 * (size_t)i];
        sx[pos] = make_float4(x[0], x[1], x[2], tag_float(i * 64 + e->h_elem[i]));
    }
    for (int32_t c = 0; c < nchunk; c++) {
        float *bx = &cbox[8 * (size_t)c];
        for (int d = 0; d < 3; d++) {
            bx[d] = 2.0f;
            bx[3 + d] = -1.0f;
        }
    }
    for (int32_t i = 0; i < n; i++) {
        float *bx = &cbox[8 * (size_t)(e->h_spos[i] / kChunkAtoms)];
        for (int d = 0; d < 3; d++) {
            bx[d] = std::min(bx[d], wf[3 * (size_t)i + d]);
            bx[3 + d] = std::max(bx[3 + d], wf[3 * (size_t)i + d]);
        }
    }
    if (!e->sx) {
        FRMC_CHECK_HIP(dalloc(e, &e->sx, sx.size()));
        FRMC_CHECK_HIP(dalloc(e, &e->cbox, cbox.size()));
        FRMC_CHECK_HIP(dalloc(e, &e->spos, (size_t)n));
    }
    e->nchunk = nchunk;
    FRMC_CHECK_HIP(hipMemcpyAsync(e->sx, sx.data(), sizeof(float4) * sx.size(), hipMemcpyHostToDevice, e->st));
    FRMC_CHECK_HIP(hipMemcpyAsync(e->cbox, cbox.data(), sizeof(float) * cbox.size(), hipMemcpyHostToDevice, e->st));
    FRMC_CHECK_HIP(
        hipMemcpyAsync(e->spos, e->h_spos.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, e->st));
    FRMC_CHECK_HIP(hipStreamSynchronize(e->st));
    e->cells_stale = false;
    e->commits_since_cells = 0;
    if (e->pend.valid) e->pend.slot = e->h_spos[e->pend.atom];
    return FRMC_OK;
}

// chunk boxes only grow as atoms move: made tight again after this many
// accepted moves
static int64_t cells_refresh_after(const frmc_engine *e) { return std::max<int64_t>(64, e->n / 4); }

static int update_cells(frmc_engine *e) {
    if (!e->cells) return FRMC_OK;
    if (e->cells_stale || e->commits_since_cells > cells_refresh_after(e)) return build_cells(e);
    return FRMC_OK;
}

void drop_cells(frmc_engine *e) {
    if (e->cells) e->cells_stale = true;
}

static void free_engine(frmc_engine *e) {
    if (!e) return;
    if (e->st) (void)hipStreamSynchronize(e->st);
    for (auto &p : e->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (hipEvent_t ev : e->event_pool) (void)hipEventDestroy(ev);
    for (void *p : e->allocs) (void)hipFree(p);
    for (void *p : e->host_allocs) (void)hipHostFree(p);
    if (e->own_stream && e->st) (void)hipStreamDestroy(e->st);
    delete e;
}

static double now_s() {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

// Wait for the results of proposals [0,k) of sequence e->seq: every word of
// the host-mapped result block carries the window's number.  A spin on pinned
// memory returns within a microsecond of k_eval_b's stores; after 60 s it
// synchronises the stream to surface a device error.
// NumPy's pairwise combine of an S(Q) constraint's leaf sums (postfix
// program over the leaves, float32 adds)
static float combine_leaves(const std::vector<int32_t> &prog, const float *leafv) {
    float stk[64];
    int sp = 0;
    for (int32_t op : prog) {
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

static int wait_done(frmc_engine *e, int32_t k) {
    volatile unsigned long long *res = e->h_res;
    const uint32_t s = e->seq;
    const int32_t W = e->resW;
    double t0 = 0.0;
    uint64_t spins = 0;
    auto wait_word = [&](size_t q) -> int {
        while ((uint32_t)(res[q] >> 32) != s) {
            if ((++spins & 0xffff) == 0) {
                if (t0 == 0.0) t0 = now_s();
                else if (now_s() - t0 > 60.0) {
                    FRMC_CHECK_HIP(hipStreamSynchronize(e->st));
                    if ((uint32_t)(res[q] >> 32) != s) {
                        set_error("window evaluation did not complete");
                        return FRMC_EHIP;
                    }
                }
            }
        }
        return FRMC_OK;
    };
    auto word = [&](size_t q) {
        const uint32_t bits = (uint32_t)res[q];
        float v;
        memcpy(&v, &bits, 4);
        return v;
    };
    if (!e->leaf_window) {
        // k_eval_b published the totals and chi-squares
        for (int32_t p = 0; p < k; p++)
            for (int32_t q = 0; q <= e->C; q++)
                if (int rc = wait_word((size_t)p * W + q)) return rc;
        __sync_synchronize();
        for (int32_t p = 0; p < k; p++) {
            e->h_tse[p] = word((size_t)p * W);
            for (int ci = 0; ci < e->C; ci++) e->h_chi2[(size_t)p * e->C + ci] = word((size_t)p * W + 1 + ci);
        }
        return FRMC_OK;
    }
    // the r-space chi-squares and the S(Q) leaf sums; each S(Q) chi-square is
    // their pairwise combine, the total np.sum([chi2_c]) from 0
    // (Engine.py:3026-3029, StructureFactorConstraints.py:742-770)
    float leafv[4096];
    for (int32_t p = 0; p < k; p++) {
        float tse = 0.0f;
        for (int ci = 0; ci < e->C; ci++) {
            const DevC &c = e->c[ci];
            float chi;
            if (c.rspace) {
                const size_t q = (size_t)p * W + 1 + ci;
                if (int rc = wait_word(q)) return rc;
                __sync_synchronize();
                chi = word(q);
            } else {
                for (int32_t l = 0; l < c.nleaves; l++)
                    if (int rc = wait_word((size_t)p * W + c.leafOff + l)) return rc;
                __sync_synchronize();
                for (int32_t l = 0; l < c.nleaves; l++) leafv[l] = word((size_t)p * W + c.leafOff + l);
                chi = combine_leaves(e->h_prog[ci], leafv);
            }
            e->h_chi2[(size_t)p * e->C + ci] = chi;
            tse = tse + chi;
        }
        e->h_tse[p] = tse;
    }
    return FRMC_OK;
}

// The decision of the last device-decided window must match the host's: k_commit
// has completed once a later window's flags are set (same stream).
static int check_device_decision(frmc_engine *e) {
    if (e->pending_check == -2) return FRMC_OK;
    const unsigned long long w = ((volatile unsigned long long *)e->h_commit)[e->pending_seq & 1];
    const int32_t dp = (int32_t)(uint32_t)w;
    const uint32_t dseq = (uint32_t)(w >> 32);
    if (dseq != e->pending_seq || dp != e->pending_check) {
        set_error("device acceptance decision (proposal %d, window %u) differs from the host's (%d, %u)", dp,
                  (unsigned)dseq, e->pending_check, (unsigned)e->pending_seq);
        e->pending_check = -2;
        return FRMC_ESTATE;
    }
    e->pending_check = -2;
    return FRMC_OK;
}

// ------------------------------------------------ k_eval_a shape tuning
//
// A shape of k_eval_a is (bins per thread per G(r) pass, proposals per
// block); every shape gives the same results.  The best shape depends on the
// window size k: few proposals want one chain per block spread over the
// chip, many want blocks whose waves share the matrix columns.  For each
// window-size bucket (ceil(log2 k), k up to kmax) every shape is timed on the
// engine stream, and the fastest is kept per shape class (device, work
// units, bins, pair types), so engines of one class tune once.
// Work units of the current layout: an r-space constraint is one unit; an
// S(Q) constraint has one per leaf of its pairwise sum (leaf layout: two
// waves each, the leaf sums go to the host) or one per 64 Q points.
static void set_units(frmc_engine *e, bool leaf) {
    e->leaf_units = leaf;
    int32_t units = 0;
    for (int ci = 0; ci < e->C; ci++) {
        e->c[ci].uOff = units;
        units += e->c[ci].rspace ? 1 : (leaf ? e->c[ci].nleaves : (e->c[ci].M + kQC - 1) / kQC);
    }
    e->nunits = units;
}

// The window's evaluation, shape s: bins per thread per G(r) pass 8 (s & 4,
// above 1,024 bins) or 4; units aligned to the XCDs (s & 8); leaf layout
// (s & 16: k_eval_a alone, the host combines the leaf sums) or 64-Q units
// (k_eval_a, then k_eval_b).  Every shape gives the same results; the best
// one depends on the window size k.  For each window-size bucket
// (ceil(log2 k), k up to kmax) every shape's evaluation is timed on the
// engine stream, twice, and the fastest is kept per shape class (device,
// constraint layout, pair types, largest window), so engines of one class
// tune once.
// the shapes the tuner times: U = 4 first, then U = 8 (above 1,024 bins),
// then the split-chain shapes (when their LDS fits): U = 2, then U = 4
// (above 2,048 bins)
static const int32_t kEvalShapes[16] = {0, 8, 16, 24, 32, 40, 48, 56, 4, 12, 20, 28, 36, 44, 52, 60};
static const int32_t kPcShapes[4] = {80, 88, 84, 92};
constexpr size_t kMaxLdsBytes = 160 * 1024;  // LDS per workgroup on gfx950

static bool pc_fits(frmc_engine *e) {
    if (e->pc_fit < 0) e->pc_fit = pc_lds_bytes(make_view(e, e->cur_stage)) <= kMaxLdsBytes ? 1 : 0;
    return e->pc_fit == 1;
}

// shape i of the tuner's list for this engine, -1 past its end
static int32_t tuner_shape(frmc_engine *e, int32_t i) {
    const int32_t nbase = e->Hmax > 4 * kEvalThreads ? 16 : 8;
    if (i < nbase) return kEvalShapes[i];
    const int32_t npc = pc_fits(e) ? (e->Hmax > 2048 ? 4 : 2) : 0;
    return i - nbase < npc ? kPcShapes[i - nbase] : -1;
}
static std::mutex g_eval_mu;
static std::map<std::tuple<int, int32_t, int32_t, int32_t, int32_t, int32_t>, std::vector<int8_t>> g_eval_tuned;
static int g_eval_forced = -1;  // frmc_set_eval_shape

static int32_t k_bucket(int32_t k) {
    int32_t b = 0;
    while (b < 15 && (1 << b) < k) b++;
    return b;
}

// one timed evaluation of shape sh for k proposals (no effect on the engine
// state: results are written with window number 0, which no window uses, and
// no delta buffer is zeroed)
static int launch_eval_shape(frmc_engine *e, int32_t k, int32_t sh) {
    set_units(e, (sh & 16) != 0);
    View V = make_view(e, e->cur_stage);
    V.seq = 0;
    V.zero_len = 0;
    V.leafmode = (sh & 16) ? 1 : 0;
    int rc = launch_eval(V, k, e->st, sh);
    if (rc == FRMC_OK && !(sh & 16)) rc = launch_eval_b(V, k, e->st);
    return rc;
}

static int tune_eval(frmc_engine *e) {
    const bool was_leaf = e->leaf_units;
    const int32_t nb = k_bucket(e->kmax) + 1;
    int32_t ns = 0;
    while (tuner_shape(e, ns) >= 0) ns++;
    constexpr int reps = 5, rounds = 2;
    std::vector<hipEvent_t> ev(2 * rounds * nb * ns, nullptr);
    int rc = FRMC_OK;
    for (auto &x : ev)
        if (rc == FRMC_OK && hipEventCreate(&x) != hipSuccess) rc = FRMC_EHIP;
    for (int32_t b = 0; b < nb && rc == FRMC_OK; b++) {
        const int32_t k = std::min<int32_t>(1 << b, e->kmax);
        for (int round = 0; round < rounds && rc == FRMC_OK; round++)
            for (int32_t i = 0; i < ns && rc == FRMC_OK; i++) {
                const size_t q = 2 * (((size_t)b * rounds + round) * ns + i);
                rc = launch_eval_shape(e, k, tuner_shape(e, i));  // warm
                if (rc == FRMC_OK && hipEventRecord(ev[q], e->st) != hipSuccess) rc = FRMC_EHIP;
                for (int r = 0; r < reps && rc == FRMC_OK; r++) rc = launch_eval_shape(e, k, tuner_shape(e, i));
                if (rc == FRMC_OK && hipEventRecord(ev[q + 1], e->st) != hipSuccess) rc = FRMC_EHIP;
            }
    }
    if (rc == FRMC_OK && hipStreamSynchronize(e->st) != hipSuccess) rc = FRMC_EHIP;
    for (int32_t b = 0; b < nb && rc == FRMC_OK; b++) {
        float best = 0.f;
        int32_t pick = 0;
        for (int32_t i = 0; i < ns && rc == FRMC_OK; i++)
            for (int round = 0; round < rounds; round++) {
                const size_t q = 2 * (((size_t)b * rounds + round) * ns + i);
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, ev[q], ev[q + 1]) != hipSuccess) {
                    rc = FRMC_EHIP;
                    break;
                }
                if ((i == 0 && round == 0) || ms < best) {
                    best = ms;
                    pick = tuner_shape(e, i);
                }
            }
        e->eval_shape[b] = (int8_t)pick;
    }
    for (int32_t b = nb; b < 16; b++) e->eval_shape[b] = e->eval_shape[nb - 1];
    for (auto x : ev)
        if (x) (void)hipEventDestroy(x);
    set_units(e, was_leaf);
    if (rc) set_error("timing the window evaluation shapes failed");
    return rc;
}

// the evaluation shape for a window of k proposals (tuned on first use)
int32_t eval_shape(frmc_engine *e, int32_t k) {
    if (g_eval_forced >= 0) return g_eval_forced;
    if (!e->eval_tuned) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        int32_t sq_points = 0;
        for (int ci = 0; ci < e->C; ci++) sq_points = sq_points * 7919 + (e->c[ci].rspace ? -e->c[ci].H : e->c[ci].M);
        const auto key = std::make_tuple(dev, e->C, e->Hmax, sq_points, e->T, e->kmax);
        bool have = false;
        {
            std::lock_guard<std::mutex> lk(g_eval_mu);
            auto it = g_eval_tuned.find(key);
            if (it != g_eval_tuned.end()) {
                for (int b = 0; b < 16; b++) e->eval_shape[b] = it->second[b];
                have = true;
            }
        }
        if (!have) {
            for (int b = 0; b < 16; b++) e->eval_shape[b] = 0;
            if (tune_eval(e) == FRMC_OK) {
                std::lock_guard<std::mutex> lk(g_eval_mu);
                g_eval_tuned[key] = std::vector<int8_t>(e->eval_shape, e->eval_shape + 16);
            }
        }
        e->eval_tuned = true;
    }
    return e->eval_shape[std::min<int32_t>(k_bucket(k), 15)];
}

// launch delta + eval for the k proposals in staging buffer s (and, with C,
// the commit of the proposal the device accepts); finish_window waits for
// the totals
static int launch_window(frmc_engine *e, int s, int32_t k, const Inline *I, const CommitCtl *C = nullptr) {
    const bool with_delta = I != nullptr;
    // the evaluation shape (tuned on the first window) fixes the S(Q) units;
    // a device-decided commit reads the totals k_eval_b leaves on the device,
    // so those windows take 64-Q units
    int32_t shape = eval_shape(e, k);
    if (C) shape &= ~(16 | 64);  // the 64-Q layout, with k_eval_b
    if ((shape & 64) && !pc_fits(e)) {
        set_error("eval shape %d needs %zu bytes of LDS per workgroup (at most %zu)", shape,
                  pc_lds_bytes(make_view(e, s)), kMaxLdsBytes);
        return FRMC_EINVAL;
    }
    set_units(e, (shape & 16) != 0);
    e->seq++;
    View V = make_view(e, s);
    // leaf layout: k_eval_a publishes the chi-squares and S(Q) leaf sums, the
    // host combines them, and k_eval_b is not needed
    V.leafmode = (shape & 16) ? 1 : 0;
    e->leaf_window = V.leafmode != 0;
    int rc;
    if (with_delta) {
        KTimer t(e, FRMC_K_DELTA);
        rc = launch_delta(V, *I, e->pend, k, e->st);
        if (rc) return rc;
        e->pend.valid = 0;  // applied by this window's k_delta
    }
    {
        KTimer t(e, FRMC_K_EVAL);
        rc = launch_eval(V, k, e->st, shape);
    }
    if (rc) return rc;
    if (!V.leafmode) {
        KTimer t(e, FRMC_K_EVAL_B);
        rc = launch_eval_b(V, k, e->st);
    }
    if (rc) return rc;
    if (C) {
        KTimer t(e, FRMC_K_COMMIT);
        rc = launch_commit(V, *I, *C, e->st);
        if (rc) return rc;
    }
    // k_eval_a (host-combined windows) or k_eval_b zeroes the other buffer
    e->dn_used[1 - e->cur_dn] = 0;
    e->dn_used[e->cur_dn] = with_delta ? k : 0;
    e->cur_dn = 1 - e->cur_dn;
    return FRMC_OK;
}

static int finish_window(frmc_engine *e, int32_t k) {
    int rc = wait_done(e, k);
    if (rc) return rc;
    rc = check_device_decision(e);
    if (rc) return rc;
    if (e->profile) collect_times(e);
    return FRMC_OK;
}

static int run_window(frmc_engine *e, int s, int32_t k, const Inline *I, const CommitCtl *C = nullptr) {
    int rc = launch_window(e, s, k, I, C);
    if (rc) return rc;
    return finish_window(e, k);
}

// Stage k proposals (group ids + new box coordinates; old ones from the host
// mirror) into the staging buffer not read by an in-flight commit, and
// launch their evaluation.  With u (acceptance draws) the device also
// commits the proposal the reference's rule accepts, queued behind the
// evaluation.  complete_eval waits for the totals.
// the host half of staging: group ids, old (host mirror) and new box
// coordinates into the host block of the staging buffer not read by an
// in-flight commit, which becomes the current one
static int stage_host(frmc_engine *e, int32_t k, const int32_t *group_ids, const float *new_box, int &s) {
    s = 1 - e->cur_stage;
    Stage &S = e->stage[s];
    S.lists = false;
    int32_t off = 0;
    for (int32_t p = 0; p < k; p++) {
        const int32_t gi = group_ids[p];
        FRMC_REQUIRE(gi >= 0 && gi < e->G, "proposal %d: group %d out of range [0,%d)", p, gi, e->G);
        S.pgroup[p] = gi;
        S.poff[p] = off;
        const int32_t g0 = e->h_goff[gi], g = e->h_goff[gi + 1] - g0;
        for (int32_t a = 0; a < g; a++) {
            const size_t id = (size_t)e->h_gatoms[g0 + a];
            for (int d = 0; d < 3; d++) S.pold[3 * (size_t)(off + a) + d] = e->h_box[3 * id + d];
        }
        off += g;
    }
    S.poff[k] = off;
    memcpy(S.pnew, new_box, sizeof(float) * 3 * (size_t)off);
    e->cur_stage = s;
    return FRMC_OK;
}

// Staging for proposals that name their moved atoms (counts[p] atoms each,
// listed in order in atoms): moves whose atoms are not one group, such as a
// swap (the group and its drawn partners).
static int stage_host_atoms(frmc_engine *e, int32_t k, const int32_t *counts, const int32_t *atoms,
                            const float *new_box, int &s) {
    s = 1 - e->cur_stage;
    Stage &S = e->stage[s];
    int32_t off = 0;
    for (int32_t p = 0; p < k; p++) {
        const int32_t g = counts[p];
        FRMC_REQUIRE(g >= 1 && g <= e->lcap, "proposal %d: %d atoms (at most %d)", p, g, e->lcap);
        S.pgroup[p] = -1;
        S.poff[p] = off;
        for (int32_t a = 0; a < g; a++) {
            const int32_t id = atoms[off + a];
            FRMC_REQUIRE(id >= 0 && id < e->n, "proposal %d: atom %d out of range [0,%d)", p, id, e->n);
            for (int32_t b = 0; b < a; b++)
                FRMC_REQUIRE(atoms[off + b] != id, "proposal %d: atom %d listed twice", p, id);
            S.patoms[off + a] = id;
            for (int d = 0; d < 3; d++) S.pold[3 * (size_t)(off + a) + d] = e->h_box[3 * (size_t)id + d];
        }
        off += g;
    }
    S.poff[k] = off;
    memcpy(S.pnew, new_box, sizeof(float) * 3 * (size_t)off);
    S.lists = true;
    e->cur_stage = s;
    return FRMC_OK;
}

static int stage_and_launch(frmc_engine *e, int32_t k, const int32_t *group_ids, const float *new_box,
                            const float *u = nullptr, float tol = 0.0f) {
    if (int rc0 = update_cells(e)) return rc0;
    int s = 0;
    if (int rc0 = stage_host(e, k, group_ids, new_box, s)) return rc0;
    Stage &S = e->stage[s];
    Inline &I = e->last_inline;
    I.k = 0;
    const bool inl = e->gmax == 1 && k <= kInline;
    if (!inl) {
        // the kernels read the device copy; the copy is ordered before them on
        // the stream.  Single-atom windows carry everything in the kernel
        // arguments and skip it (the host block still backs note_commit).
        FRMC_CHECK_HIP(hipMemcpyAsync(S.dev, S.host, S.bytes, hipMemcpyHostToDevice, e->st));
    }
    if (inl) {
        I.k = k;
        for (int32_t p = 0; p < k; p++) {
            const int32_t id = e->h_gatoms[e->h_goff[group_ids[p]]];
            I.atom[p] = id;
            I.el[p] = e->h_elem[id];
            for (int d = 0; d < 3; d++) {
                I.old[p][d] = e->h_box[3 * (size_t)id + d];
                I.nw[p][d] = new_box[3 * (size_t)p + d];
            }
        }
    }
    int rc;
    e->win_dn = e->dnbuf[e->cur_dn];  // the buffer this window's deltas go to
    if (u && e->gmax == 1) {
        // single-atom moves of the native loop: the host's decision is
        // committed by the next window's k_delta (or flush_commit)
        rc = launch_window(e, s, k, &I);
        e->device_decided = false;
        e->deferred = true;
    } else if (u && k <= kInline) {
        e->deferred = false;
        CommitCtl C;
        C.k = k;
        C.p = -1;
        C.cur = e->cur_tse;
        C.tol = tol;
        C.dn = e->dnbuf[e->cur_dn];  // the buffer this window's deltas go to
        for (int32_t p = 0; p < k; p++) C.u[p] = u[p];
        rc = launch_window(e, s, k, &I, &C);
        e->device_decided = true;
    } else {
        e->deferred = false;
        rc = launch_window(e, s, k, &I);
        e->device_decided = false;
    }
    return rc;
}

// A single-atom window of the native loop for a batched launch: the same
// staging, window bookkeeping and deferred commit as stage_and_launch, with
// the proposals written to props and the window to W instead of launched.
int batch_prepare(frmc_engine *e, int32_t k, const int32_t *group_ids, const float *new_box, BatchWin &W,
                  Prop *props) {
    FRMC_REQUIRE(e->gmax == 1 && k >= 1 && k <= e->kmax, "batched windows need single-atom groups and 1 <= k <= %d",
                 e->kmax);
    int s = 0;
    if (int rc = stage_host(e, k, group_ids, new_box, s)) return rc;
    for (int32_t p = 0; p < k; p++) {
        const int32_t id = e->h_gatoms[e->h_goff[group_ids[p]]];
        Prop &q = props[p];
        q.atom = id;
        q.el = e->h_elem[id];
        for (int d = 0; d < 3; d++) {
            q.old[d] = e->h_box[3 * (size_t)id + d];
            q.nw[d] = new_box[3 * (size_t)p + d];
        }
    }
    e->last_inline.k = 0;
    e->win_dn = e->dnbuf[e->cur_dn];
    // launch_window's bookkeeping
    e->seq++;
    W.k = k;
    W.seq = e->seq;
    W.dn = e->dnbuf[e->cur_dn];
    W.dn_zero = e->dnbuf[1 - e->cur_dn];
    W.zero_len = (int64_t)e->dn_used[1 - e->cur_dn] * e->dnStride;
    W.P = e->pend;
    e->pend.valid = 0;
    e->dn_used[1 - e->cur_dn] = 0;
    e->dn_used[e->cur_dn] = k;
    e->cur_dn = 1 - e->cur_dn;
    e->device_decided = false;
    e->deferred = true;
    e->leaf_window = false;  // k_eval_b_batch publishes the totals
    return FRMC_OK;
}

void use_unit_layout(frmc_engine *e, bool leaf) {
    set_units(e, leaf);
}

// Apply an accepted move still left for a next window's k_delta, as a
// k_commit of its own (before anything else reads or replaces the state), on
// the engine's stream or on stream `on` (the caller then orders the engine's
// stream after it).
int flush_commit(frmc_engine *e, hipStream_t on) {
    if (!e->pend.valid) return FRMC_OK;
    Inline I1;
    I1.k = 1;
    I1.atom[0] = e->pend.atom;
    I1.el[0] = e->pend.el;
    for (int d = 0; d < 3; d++) {
        I1.old[0][d] = e->pend.old[d];
        I1.nw[0][d] = e->pend.nw[d];
    }
    CommitCtl C;
    C.k = 1;
    C.p = 0;
    C.cur = e->cur_tse;
    C.tol = 0.0f;
    C.dn = e->pend.dn;
    View V = make_view(e, e->cur_stage);
    e->pend.valid = 0;
    if (on && on != e->st) return launch_commit(V, I1, C, on);
    KTimer t(e, FRMC_K_COMMIT);
    return launch_commit(V, I1, C, e->st);
}

static int complete_eval(frmc_engine *e, int32_t k) {
    int rc = finish_window(e, k);
    if (rc) return rc;
    e->last_k = k;
    return FRMC_OK;
}

static int stage_and_eval(frmc_engine *e, int32_t k, const int32_t *group_ids, const float *new_box,
                          const float *u = nullptr, float tol = 0.0f) {
    int rc = stage_and_launch(e, k, group_ids, new_box, u, tol);
    if (rc) return rc;
    return complete_eval(e, k);
}

// host mirror of a committed proposal p of the last window (the group's new
// box coordinates, the totals)
static void note_commit(frmc_engine *e, int32_t p) {
    const Stage &S = e->stage[e->cur_stage];
    const int32_t gi = S.pgroup[p], a0 = S.poff[p];
    const int32_t g0 = S.lists ? 0 : e->h_goff[gi], g = S.lists ? S.poff[p + 1] - a0 : e->h_goff[gi + 1] - g0;
    for (int32_t a = 0; a < g; a++) {
        const size_t id = (size_t)(S.lists ? S.patoms[a0 + a] : e->h_gatoms[g0 + a]);
        for (int d = 0; d < 3; d++) {
            const float x = S.pnew[3 * (size_t)(a0 + a) + d];
            e->h_box[3 * id + d] = x;
            if (!(fabsf(x) <= kCellsMaxCoord)) e->cells_stale = true;  // the rebuild turns the copy off
        }
    }
    e->commits_since_cells++;
    e->cur_tse = e->h_tse[p];
    for (int ci = 0; ci < e->C; ci++) e->cur_chi2[ci] = e->h_chi2[(int64_t)p * e->C + ci];
    e->last_k = 0;  // the window now describes a stale state
}

// A window whose commit the device decides, in two halves so that several
// engines can have windows in flight: launch_device_decided stages and
// launches it, finish_device_decided waits and returns the totals.  The host
// applies the same rule to them, then calls settle_window.
int launch_device_decided(frmc_engine *e, int32_t k, const int32_t *group_ids, const float *new_box,
                          const float *u, float tol) {
    return stage_and_launch(e, k, group_ids, new_box, u, tol);
}

int finish_device_decided(frmc_engine *e, int32_t k, float *tse_out) {
    int rc = complete_eval(e, k);
    if (rc) return rc;
    memcpy(tse_out, e->h_tse.data(), sizeof(float) * k);
    return FRMC_OK;
}

// commit what is pending, wait for the stream and verify the last
// device-made decision
int finish_windows(frmc_engine *e) {
    int rc = flush_commit(e);
    if (rc) return rc;
    FRMC_CHECK_HIP(hipStreamSynchronize(e->st));
    return check_device_decision(e);
}

// The host's decision for the last window: p >= 0 committed (by the device
// when it decided, else by a commit launched here), -1 none.
int settle_window(frmc_engine *e, int32_t p) {
    if (e->deferred) {
        e->deferred = false;
        if (p < 0) {
            e->last_k = 0;
            return FRMC_OK;
        }
        const Stage &S = e->stage[e->cur_stage];
        const int32_t a0 = S.poff[p];
        const int32_t id = e->h_gatoms[e->h_goff[S.pgroup[p]]];
        e->pend.valid = 1;
        e->pend.atom = id;
        e->pend.el = e->h_elem[id];
        for (int d = 0; d < 3; d++) {
            e->pend.old[d] = S.pold[3 * (size_t)a0 + d];
            e->pend.nw[d] = S.pnew[3 * (size_t)a0 + d];
        }
        e->pend.dn = e->win_dn + (int64_t)p * e->dnStride;
        e->pend.slot = e->cells ? e->h_spos[id] : -1;
        note_commit(e, p);
        return FRMC_OK;
    }
    if (e->device_decided) {
        e->pending_check = p;
        e->pending_seq = e->seq;
        if (p >= 0) note_commit(e, p);
        else e->last_k = 0;
        return FRMC_OK;
    }
    if (p < 0) {
        e->last_k = 0;
        return FRMC_OK;
    }
    return frmc_engine_commit(e, p);
}

// ------------------------------------------------------ device window loop
//
// Batches of windows of one size k whose walk and draws run on the device
// (k_loop_step, window_loop.hip): per window the host only queues three or
// four launches, and it synchronises once per batch.  Between batches the
// host mirror (box and real coordinates, totals, pending commit) is brought
// up to date from the batch's accepted moves, so the host loop can take over
// at any batch boundary.

// frmc_set_window_loop: -1 default, 0 off, 1 on.  The default leaves single
// engines on the host loop: the step's latency (two memory round trips and
// a launch, about 7 us) equals the host's round trip between windows, and
// runs measured 2-6 % slower with the device loop (200,000-atom SiO2:
// 117.3k against 122.1k moves/s; 13,500-atom B2: 159.6k against 170.3k).
static int g_loop_mode = -1;

// The S(Q) leaf combines (NumPy's pairwise sum over the leaf sums, a postfix
// program of float32 adds) as trees for the device loop: every add is a node
// whose value goes to a row word of its own after the result words, and the
// nodes are grouped by level (one above the deeper child; leaves at 0), so
// that k_loop_step runs a level's adds in parallel.  Built once per engine;
// false when the trees exceed the step's tables (the host loop runs then).
static bool loop_trees(frmc_engine *e) {
    if (e->loop_nnodes >= 0) return true;
    if (e->loop_nnodes == -2) return false;
    struct Node {
        int32_t x, y, z, lev;
    };
    std::vector<Node> all;
    int32_t next = e->resW;
    e->loop_nnodes = -2;
    for (int ci = 0; ci < e->C; ci++) {
        if (e->c[ci].rspace) {
            e->loop_root[ci] = 1 + ci;
            continue;
        }
        std::vector<std::pair<int32_t, int32_t>> stk;  // (row word, level)
        for (int32_t op : e->h_prog[ci]) {
            if (op >= 0) {
                stk.push_back({e->c[ci].leafOff + op, 0});
            } else {
                if (stk.size() < 2) return false;
                const auto b = stk.back();
                stk.pop_back();
                const auto x = stk.back();
                stk.pop_back();
                const int32_t lev = std::max(x.second, b.second) + 1;
                all.push_back({x.first, b.first, next, lev});
                stk.push_back({next++, lev});
            }
        }
        if (stk.size() != 1) return false;
        e->loop_root[ci] = stk[0].first;
    }
    int32_t maxlev = 0;
    for (const Node &n : all) maxlev = std::max(maxlev, n.lev);
    if ((int32_t)all.size() > kLoopNodes || maxlev > kLoopLevels || e->resW > kLoopResW) return false;
    std::stable_sort(all.begin(), all.end(), [](const Node &u, const Node &v) { return u.lev < v.lev; });
    std::vector<int4> h(all.size());
    for (int32_t l = 0; l <= maxlev; l++) e->loop_lev_off[l] = 0;
    for (size_t i = 0; i < all.size(); i++) {
        h[i] = make_int4(all[i].x, all[i].y, all[i].z, all[i].lev);
        for (int32_t l = all[i].lev; l <= maxlev; l++) e->loop_lev_off[l] = (int32_t)i + 1;
    }
    if (!all.empty()) {
        if (dupload(e, &e->loop_nodes, h.data(), h.size()) != hipSuccess) return false;
    }
    e->loop_nlev = maxlev;
    e->loop_rs = next;
    e->loop_nnodes = (int32_t)all.size();
    return true;
}

bool loop_usable(frmc_engine *e, int32_t kmax_run) {
    if (g_loop_mode != 1 || e->gmax != 1 || !e->have_real || kmax_run > kLoopMaxK) return false;
    return loop_trees(e);
}

// Batched replicas take the loop by default: there the host's walks and
// draws of all replicas, not the GPU, bound a step.
bool loop_batch_usable(frmc_engine *e, int32_t kmax_run) {
    return g_loop_mode != 0 && e->gmax == 1 && e->have_real && kmax_run <= kLoopMaxK && kmax_run <= e->kmax &&
           e->resW <= kLoopResW;
}

// The loop's buffers (first use; single: also the single-engine state, logs
// and proposal table) and the device copy of the real coordinates (queued
// on the engine stream).
int loop_alloc(frmc_engine *e, bool single) {
    if (!e->d_real) {
        FRMC_CHECK_HIP(dalloc(e, &e->d_real, 3 * (size_t)e->n));
        FRMC_CHECK_HIP(dalloc(e, &e->d_res, (size_t)e->kmax * e->resW));
    }
    if (single && !e->loop) {
        const size_t cap = (size_t)kLoopWindows * kLoopMaxK;
        FRMC_CHECK_HIP(dalloc(e, &e->loop, 1));
        FRMC_CHECK_HIP(halloc(e, &e->h_loop, 1));
        FRMC_CHECK_HIP(dalloc(e, &e->loop_glog, cap));
        FRMC_CHECK_HIP(halloc(e, &e->h_loop_glog, cap));
        FRMC_CHECK_HIP(dalloc(e, &e->loop_alog, cap));
        FRMC_CHECK_HIP(halloc(e, &e->h_loop_alog, cap));
        FRMC_CHECK_HIP(dalloc(e, &e->loop_tlog, cap));
        FRMC_CHECK_HIP(halloc(e, &e->h_loop_tlog, cap));
        FRMC_CHECK_HIP(dalloc(e, &e->loop_acc, (size_t)kLoopWindows));
        FRMC_CHECK_HIP(halloc(e, &e->h_loop_acc, (size_t)kLoopWindows));
        FRMC_CHECK_HIP(dalloc(e, &e->loop_props, (size_t)kLoopMaxK));
    }
    if (e->groups_ident < 0) {
        bool id = e->G == e->n;
        for (int32_t g = 0; g < e->G && id; g++) id = e->h_goff[g] == g && e->h_gatoms[g] == g;
        e->groups_ident = id ? 1 : 0;
    }
    FRMC_CHECK_HIP(hipMemcpyAsync(e->d_real, e->h_real.data(), sizeof(float) * 3 * (size_t)e->n,
                                  hipMemcpyHostToDevice, e->st));
    return FRMC_OK;
}

// a run's starting loop state: both streams, the totals, the pending commit
void loop_init_state(frmc_engine *e, LoopState &H, const frmc_mt19937 &py, const frmc_mt19937 &np) {
    memcpy(H.py.mt, py.mt, sizeof(H.py.mt));
    H.py.pos = py.pos;
    memcpy(H.np.mt, np.mt, sizeof(H.np.mt));
    H.np.pos = np.pos;
    H.h.cur_tse = e->cur_tse;
    for (int ci = 0; ci < kMaxC; ci++) H.h.cur_chi2[ci] = e->cur_chi2[ci];
    H.h.pad = 0;
    H.h.step = 0;
    H.h.nacc = H.h.ntol = H.h.error = H.h.kw = 0;
    H.h.pend = e->pend;
}

// Allocate the loop's buffers (first use) and load its state: both streams,
// the totals, the device copy of the real coordinates.
int loop_begin(frmc_engine *e, const frmc_mt19937 &py, const frmc_mt19937 &np) {
    // nothing queued may still read the pinned block
    FRMC_CHECK_HIP(hipStreamSynchronize(e->st));
    if (int rc = loop_alloc(e, true)) return rc;
    loop_init_state(e, *e->h_loop, py, np);
    FRMC_CHECK_HIP(hipMemcpyAsync(e->loop, e->h_loop, offsetof(LoopState, gid), hipMemcpyHostToDevice, e->st));
    e->loop_maxabs = 0.0f;
    for (float x : e->h_box) e->loop_maxabs = std::max(e->loop_maxabs, fabsf(x));
    FRMC_CHECK_HIP(hipStreamSynchronize(e->st));  // the pinned block is rewritten per batch
    return FRMC_OK;
}

// The host mirror after a batch of the loop: the accepted moves' box and
// real coordinates, the totals, the pending commit; an error the step
// reported.
int loop_apply(frmc_engine *e, const LoopState &H, const LoopAcc *acc, int32_t M, int64_t nlog) {
    if (H.h.error) {
        set_error(H.h.error == kLoopErrDirection ? "a direction draw of length 0 (the device window loop does not redraw)"
                                                 : "a window's draws ran past four MT19937 state blocks");
        return FRMC_ESTATE;
    }
    FRMC_REQUIRE(H.h.nacc >= 0 && H.h.nacc <= M && H.h.step >= 0 && H.h.step <= nlog, "bad loop batch state");
    for (int32_t i = 0; i < H.h.nacc; i++) {
        const LoopAcc &a = acc[i];
        FRMC_REQUIRE(a.atom >= 0 && a.atom < e->n, "bad accepted move");
        for (int d = 0; d < 3; d++) {
            e->h_box[3 * (size_t)a.atom + d] = a.box[d];
            e->h_real[3 * (size_t)a.atom + d] = a.real[d];
            e->loop_maxabs = std::max(e->loop_maxabs, fabsf(a.box[d]));
            if (!(fabsf(a.box[d]) <= kCellsMaxCoord)) e->cells_stale = true;
        }
    }
    e->commits_since_cells += H.h.nacc;
    e->cur_tse = H.h.cur_tse;
    for (int ci = 0; ci < e->C; ci++) e->cur_chi2[ci] = H.h.cur_chi2[ci];
    e->pend = H.h.pend;
    e->win_dn = H.h.pend.dn;
    e->deferred = false;
    e->device_decided = false;
    e->last_k = 0;
    return FRMC_OK;
}

// Windows the next batch may hold (at most want): the chunk walk's sorted
// copy is brought up to date first, and a batch holds no more accepted moves
// than the copy takes before its next rebuild, nor enough to carry a box
// coordinate past the magnitude the chunk test allows (each move changes a
// box coordinate by at most maxAmp times the norm of a column of the
// reciprocal basis).
int loop_plan(frmc_engine *e, float maxAmp, int32_t want, int32_t &M) {
    if (int rc = update_cells(e)) return rc;
    int64_t m = want;
    if (e->cells && !e->cells_stale) {
        m = std::min<int64_t>(m, std::max<int64_t>(1, cells_refresh_after(e) + 1 - e->commits_since_cells));
        double col = 0.0;
        for (int d = 0; d < 3; d++) {
            double s2 = 0.0;
            for (int i = 0; i < 3; i++) s2 += (double)e->recip[3 * i + d] * e->recip[3 * i + d];
            col = std::max(col, sqrt(s2));
        }
        const double per = (double)maxAmp * col * (1.0 + 1e-3) + 1e-6;
        const double room = (double)kCellsMaxCoord - e->loop_maxabs;
        m = std::min<int64_t>(m, room > 0.0 ? (int64_t)std::floor(room / per) : 0);
    }
    M = (int32_t)std::max<int64_t>(0, m);
    return FRMC_OK;
}

// One batch of M windows of k proposals; the host mirror is updated from
// the batch's accepted moves and e->h_loop holds the batch's state and logs.
int loop_batch(frmc_engine *e, int32_t k, int32_t M, float minAmp, float maxAmp, float tol) {
    FRMC_REQUIRE(M >= 1 && M <= kLoopWindows && k >= 1 && k <= kLoopMaxK && k <= e->kmax, "loop batch %d x %d", M,
                 k);
    const int32_t shape = eval_shape(e, k);
    if ((shape & 64) && !pc_fits(e)) {
        set_error("eval shape %d needs %zu bytes of LDS per workgroup (at most %zu)", shape,
                  pc_lds_bytes(make_view(e, 0)), kMaxLdsBytes);
        return FRMC_EINVAL;
    }
    const bool leaf = (shape & 16) != 0;
    LoopState &H = *e->h_loop;
    H.h.pend = e->pend;
    FRMC_CHECK_HIP(hipMemcpyAsync(&e->loop->h.pend, &H.h.pend, sizeof(Pending), hipMemcpyHostToDevice, e->st));
    LoopArgs A{};
    A.L = e->loop;
    A.resW = e->resW;
    A.C = e->C;
    A.dnStride = e->dnStride;
    A.tol = tol;
    A.res = e->d_res;
    A.nodes = e->loop_nodes;
    A.nnodes = e->loop_nnodes;
    A.nlev = e->loop_nlev;
    A.rs = e->loop_rs;
    for (int l = 0; l <= kLoopLevels; l++) A.lev_off[l] = e->loop_lev_off[l];
    for (int ci = 0; ci < e->C; ci++) {
        A.rspace[ci] = e->c[ci].rspace;
        A.root[ci] = e->loop_root[ci];
    }
    A.G = e->G;
    A.nbits = bit_length(e->G);
    A.ident = e->groups_ident;
    A.minAmp = minAmp;
    A.maxAmp = maxAmp;
    for (int q = 0; q < 9; q++) A.recip[q] = e->recip[q];
    A.goff = e->goff;
    A.gatoms = e->gatoms;
    A.elem = e->elem;
    A.spos = (e->cells && !e->cells_stale) ? e->spos : nullptr;
    A.xyz = e->xyz;
    A.real = e->d_real;
    A.props = e->loop_props;
    A.glog = e->loop_glog;
    A.alog = e->loop_alog;
    A.tlog = e->loop_tlog;
    A.acc = e->loop_acc;
    A.left = (int64_t)1 << 62;  // the caller's batches fit the moves left
    Inline I0;
    I0.k = 0;
    Pending P0{};
    P0.valid = 0;
    for (int32_t w = 0; w <= M; w++) {
        A.reset = w == 0;
        A.dk = w < M ? k : 0;
        if (int rc = launch_loop_step(A, e->st)) return rc;
        if (w == M) break;
        set_units(e, leaf);
        e->seq++;
        View V = make_view(e, 0);
        V.leafmode = leaf ? 1 : 0;
        V.out_res = e->d_res;
        V.pend_dev = &e->loop->h.pend;
        V.props = e->loop_props;
        e->leaf_window = leaf;
        {
            KTimer t(e, FRMC_K_DELTA);
            if (int rc = launch_delta(V, I0, P0, k, e->st)) return rc;
        }
        {
            KTimer t(e, FRMC_K_EVAL);
            if (int rc = launch_eval(V, k, e->st, shape)) return rc;
        }
        if (!leaf) {
            KTimer t(e, FRMC_K_EVAL_B);
            if (int rc = launch_eval_b(V, k, e->st)) return rc;
        }
        e->dn_used[1 - e->cur_dn] = 0;
        e->dn_used[e->cur_dn] = k;
        e->cur_dn = 1 - e->cur_dn;
        // the next step walks this window
        A.wk = k;
        A.leafmode = V.leafmode;
        A.win_dn = V.dn;
    }
    const size_t nlog = (size_t)M * k;
    FRMC_CHECK_HIP(hipMemcpyAsync(e->h_loop, e->loop, sizeof(LoopState), hipMemcpyDeviceToHost, e->st));
    FRMC_CHECK_HIP(hipMemcpyAsync(e->h_loop_glog, e->loop_glog, sizeof(int32_t) * nlog, hipMemcpyDeviceToHost, e->st));
    FRMC_CHECK_HIP(hipMemcpyAsync(e->h_loop_alog, e->loop_alog, nlog, hipMemcpyDeviceToHost, e->st));
    FRMC_CHECK_HIP(hipMemcpyAsync(e->h_loop_tlog, e->loop_tlog, sizeof(float) * nlog, hipMemcpyDeviceToHost, e->st));
    FRMC_CHECK_HIP(
        hipMemcpyAsync(e->h_loop_acc, e->loop_acc, sizeof(LoopAcc) * (size_t)M, hipMemcpyDeviceToHost, e->st));
    FRMC_CHECK_HIP(hipStreamSynchronize(e->st));
    if (e->profile) collect_times(e);
    e->loop_windows += M;
    return loop_apply(e, H, e->h_loop_acc, M, (int64_t)nlog);
}

}  // namespace frmc

using namespace frmc;

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
#define EFAIL(code, ...)        \
    do {                        \
        set_error(__VA_ARGS__); \
        rc = code;              \
        goto fail;              \
    } while (0)
#define ECHK(x)                                                                        \
    do {                                                                               \
        err = (x);                                                                     \
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
        e->ortho = e->isPBC && d->basis[1] == 0.f && d->basis[2] == 0.f && d->basis[3] == 0.f &&
                   d->basis[5] == 0.f && d->basis[6] == 0.f && d->basis[7] == 0.f;
        int dev = 0;
        ECHK(hipGetDevice(&dev));
        e->device = dev;
        ECHK(hipDeviceGetAttribute(&e->ncu, hipDeviceAttributeMultiprocessorCount, dev));
        if (d->stream) {
            e->st = (hipStream_t)d->stream;
        } else {
            ECHK(hipStreamCreate(&e->st));
            e->own_stream = true;
        }
        e->h_goff.assign(d->group_offsets, d->group_offsets + d->ngroups + 1);
        if (e->h_goff[0] != 0) EFAIL(FRMC_EINVAL, "group_offsets[0] must be 0");
        for (int32_t gi = 0; gi < d->ngroups; gi++) {
            const int32_t g = e->h_goff[gi + 1] - e->h_goff[gi];
            if (g < 1 || g > kMaxG) EFAIL(FRMC_EINVAL, "group %d has %d atoms (1..%d supported)", gi, g, kMaxG);
            e->gmax = std::max(e->gmax, g);
        }
        e->lcap = std::min(kMaxG, 2 * e->gmax);
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
        e->h_elem.assign(d->elem, d->elem + d->n);
        ECHK(dupload(e, &e->goff, d->group_offsets, (size_t)d->ngroups + 1));
        ECHK(dupload(e, &e->gatoms, d->group_atoms, (size_t)e->total_gatoms));
        e->h_box.assign(d->box, d->box + 3 * (size_t)d->n);
        if (d->real) {
            e->h_real.assign(d->real, d->real + 3 * (size_t)d->n);
            e->have_real = true;
        }
        int32_t off = 0;
        for (int ci = 0; ci < e->C; ci++) {
            const frmc_constraint_desc &cd = d->constraints[ci];
            DevC &dc = e->c[ci];
            memset(&dc, 0, sizeof(dc));
            if (cd.kind != FRMC_PDC && cd.kind != FRMC_SFC && cd.kind != FRMC_PCF && cd.kind != FRMC_RSFC)
                EFAIL(FRMC_EINVAL, "constraint %d: unknown kind", ci);
            if (cd.histSize < 1) EFAIL(FRMC_EINVAL, "constraint %d: histSize must be >= 1", ci);
            if (!(cd.bin > 0.f)) EFAIL(FRMC_EINVAL, "constraint %d: bin must be > 0", ci);
            dc.kind = cd.kind;
            dc.rspace = cd.kind == FRMC_PDC || cd.kind == FRMC_PCF;
            dc.H = cd.histSize;
            dc.M = dc.rspace ? cd.histSize : cd.nQ;
            if (dc.M < 1) EFAIL(FRMC_EINVAL, "constraint %d: no model points", ci);
            if ((int64_t)e->nel * e->nel * cd.histSize * 2 + 1 > INT32_MAX)
                EFAIL(FRMC_EINVAL, "constraint %d: nel^2*histSize too large", ci);
            dc.g = make_grid(cd.minDistance, cd.maxDistance, cd.bin, cd.histSize);
            if (ci > 0) {
                const Grid &pg = e->c[ci - 1].g;
                dc.sameGrid = pg.minD == dc.g.minD && pg.maxD == dc.g.maxD && pg.bin == dc.g.bin &&
                              pg.s_lo == dc.g.s_lo && pg.s_hi == dc.g.s_hi && pg.H == dc.g.H;
            }
            dc.hasW = cd.dataWeights != nullptr;
            dc.hasSF = cd.scaleFactor != 1.0f;
            dc.sf = cd.scaleFactor;
            if (!cd.pairWeights || !cd.pairDensity || !cd.shellVolumes || !cd.prefactor || !cd.experimental)
                EFAIL(FRMC_EINVAL, "constraint %d: NULL array", ci);
            if (!dc.rspace && !cd.sqMatrix) EFAIL(FRMC_EINVAL, "constraint %d: S(Q) needs sqMatrix", ci);
            float *wp, *Dp, *sv, *pf, *ex, *dw = nullptr, *sm = nullptr, *si = nullptr;
            ECHK(dupload(e, &wp, cd.pairWeights, (size_t)e->T));
            ECHK(dupload(e, &Dp, cd.pairDensity, (size_t)e->T));
            ECHK(dupload(e, &sv, cd.shellVolumes, (size_t)dc.H));
            ECHK(dupload(e, &pf, cd.prefactor, (size_t)dc.H));
            ECHK(dupload(e, &ex, cd.experimental, (size_t)dc.M));
            if (cd.dataWeights) ECHK(dupload(e, &dw, cd.dataWeights, (size_t)dc.M));
            if (!dc.rspace) {
                // k_eval_a streams one Q column per lane: the matrix is kept transposed,
                // each column zero-padded to a multiple of the kernel's 32-row batches
                dc.hpb = (dc.H + 31) / 32 * 32;
                std::vector<float> mt((size_t)dc.M * dc.hpb, 0.0f);
                for (int32_t r = 0; r < dc.H; r++)
                    for (int32_t q = 0; q < dc.M; q++) mt[(size_t)q * dc.hpb + r] = cd.sqMatrix[(size_t)r * dc.M + q];
                ECHK(dupload(e, &sm, mt.data(), mt.size()));
                // and interleaved by groups of four rows, [r/4][q][r%4]: a wave's
                // 64 lanes read four rows of 64 consecutive Q points as one
                // contiguous kilobyte (k_eval_a indexes it with 32-bit offsets)
                if ((int64_t)dc.hpb * dc.M >= ((int64_t)1 << 31))
                    EFAIL(FRMC_EINVAL, "constraint %d: S(Q) matrix of %d x %d too large", ci, dc.hpb, dc.M);
                std::vector<float> mi((size_t)dc.M * dc.hpb, 0.0f);
                for (int32_t r = 0; r < dc.H; r++)
                    for (int32_t q = 0; q < dc.M; q++)
                        mi[((size_t)(r / 4) * dc.M + q) * 4 + r % 4] = cd.sqMatrix[(size_t)r * dc.M + q];
                ECHK(dupload(e, &si, mi.data(), mi.size()));
            }
            dc.w = wp;
            dc.D = Dp;
            dc.shellV = sv;
            dc.pref = pf;
            dc.expv = ex;
            dc.dataW = dw;
            dc.sqm = sm;
            dc.sqi = si;
            std::vector<int32_t> lv, pg;
            pairwise_plan(dc.M, lv, pg);
            int32_t *lvp, *pgp;
            ECHK(dupload(e, &lvp, lv.data(), lv.size()));
            ECHK(dupload(e, &pgp, pg.data(), pg.size()));
            dc.leaves = lvp;
            dc.prog = pgp;
            dc.nleaves = (int32_t)(lv.size() / 2);
            dc.nprog = (int32_t)pg.size();
            e->h_prog[ci] = pg;
            ECHK(dalloc(e, &dc.base, (size_t)e->T * dc.H));
            ECHK(dalloc(e, &dc.slot, 2 * (size_t)e->nel * e->nel * dc.H + 1));
            if (dc.sameGrid) {
                dc.dnOff = e->c[ci - 1].dnOff;  // identical deltas: one block serves both constraints
            } else {
                dc.dnOff = off;
                off += e->T * dc.H;
            }
            e->Hmax = std::max(e->Hmax, dc.H);
            e->Mmax = std::max(e->Mmax, dc.M);
            e->maxLeaves = std::max(e->maxLeaves, dc.nleaves);
        }
        e->dnStride = off;
        // work units: an r-space constraint, or a leaf (at most 128 Q points,
        // two waves) of an S(Q) constraint's pairwise sum; result words per
        // proposal: total, chi-squares, then the leaf sums
        // result words per proposal: total, chi-squares, then the leaf sums
        int32_t words = 1 + e->C;
        for (int ci = 0; ci < e->C; ci++) {
            e->c[ci].leafOff = words;
            if (!e->c[ci].rspace) words += e->c[ci].nleaves;
        }
        e->resW = words;
        // the LDS checks below take the larger (leaf) layout's needs; the
        // tuned shape picks the layout per window
        set_units(e, false);
        for (int ci = 0; ci < e->C; ci++)
            if (e->c[ci].nleaves > 4096) EFAIL(FRMC_EINVAL, "too many Q points (%d) in constraint %d", e->c[ci].M, ci);
        {
            View V = make_view(e, 0);
            // the evaluation kernels keep a constraint's model array in LDS
            // (150 KB leaves room for the batched kernels' staged records)
            if (eval_lds_bytes(V) > 150 * 1024) EFAIL(FRMC_EINVAL, "model arrays need more than 150 KB of LDS");
            if (eval_b_lds_bytes(V) > 150 * 1024)
                EFAIL(FRMC_EINVAL, "an S(Q) constraint with %d Q points needs %zu bytes of LDS in the chi-square "
                      "kernel (at most 150 KB)", V.Mmax, eval_b_lds_bytes(V));
        }
        const size_t K = (size_t)e->kmax;
        for (int b = 0; b < 2; b++) {
            ECHK(dalloc(e, &e->dnbuf[b], K * e->dnStride));
            ECHK(hipMemset(e->dnbuf[b], 0, sizeof(int32_t) * K * e->dnStride));
            Stage &S = e->stage[b];
            // group ids, K+1 offsets (padded to even), the atom lists, old and
            // new box coordinates
            const size_t L = (size_t)e->lcap;
            const size_t ni = 2 * K + 2 + ((K * L + 1) & ~(size_t)1), nf = 2 * K * L * 3;
            S.bytes = 4 * (ni + nf);
            char *hb = nullptr, *db = nullptr;
            ECHK(halloc(e, &hb, S.bytes));
            ECHK(dalloc(e, &db, S.bytes));
            S.host = hb;
            S.dev = db;
            S.pgroup = (int32_t *)hb;
            S.poff = S.pgroup + K;
            S.patoms = S.poff + K + 2;
            S.pold = (float *)(hb + 4 * ni);
            S.pnew = S.pold + K * L * 3;
            S.d_pgroup = (int32_t *)db;
            S.d_poff = S.d_pgroup + K;
            S.d_patoms = S.d_poff + K + 2;
            S.d_pold = (float *)(db + 4 * ni);
            S.d_pnew = S.d_pold + K * L * 3;
        }
        ECHK(dalloc(e, &e->sqbuf, K * e->C * e->Mmax));
        ECHK(dalloc(e, &e->chi2d, K * e->C));
        // blocks of k_eval finished per proposal (the last one resets it)
        ECHK(dalloc(e, &e->cnt, K));
        ECHK(hipMemset(e->cnt, 0, sizeof(int32_t) * K));
        ECHK(halloc(e, &e->h_res, K * e->resW));
        for (size_t q = 0; q < K * (size_t)e->resW; q++) e->h_res[q] = 0;
        e->h_tse.assign(K, 0.0f);
        e->h_chi2.assign(K * e->C, 0.0f);
        ECHK(dalloc(e, &e->tse_dev, K));
        ECHK(halloc(e, &e->h_commit, 2));
        e->h_commit[0] = e->h_commit[1] = 0xffffffffull;
        e->pend.slot = -1;
        // the sorted copy for k_delta's chunk walk: periodic boxes whose groups
        // are single atoms (atom index and element share one 32-bit tag)
        e->cells = e->isPBC && e->gmax == 1 && d->n < (1 << 24) && g_cells_mode != 0 &&
                   (g_cells_mode == 1 || d->n >= kCellsMinAtoms);
        if (e->cells) {
            rc = build_cells(e);
            if (rc) goto fail;
        }
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
    if (int rc0 = flush_commit(e)) return rc0;
    hipStream_t st = e->st;
    for (int ci = 0; ci < e->C; ci++) {
        DevC &dc = e->c[ci];
        HistArgs A;
        A.g = dc.g;
        A.nel = e->nel;
        A.size = (int64_t)e->nel * e->nel * dc.H;
        int rc = launch_fill(dc.slot, 2 * A.size + 1, st);
        if (!rc) rc = launch_full(e->xyz, e->n, e->B, e->isPBC, e->ortho != 0, e->mol, e->elem, A, dc.slot, st);
        if (!rc) rc = launch_symmetrise(dc.base, dc.slot, e->nel, dc.H, st);
        if (rc) return rc;
    }
    return frmc_engine_refresh(e, chi2_out, tse_out);
}

int frmc_engine_refresh(frmc_engine *e, float *chi2_out, float *tse_out) {
    // the current state evaluated as one zero-delta window (the delta buffer
    // of the next window is all zero)
    FRMC_REQUIRE(e, "NULL engine");
    if (int rc0 = flush_commit(e)) return rc0;
    const int cur = e->cur_dn;
    int rc = run_window(e, e->cur_stage, 1, nullptr);
    if (rc) return rc;
    e->cur_dn = cur;  // the zero buffer was not written; keep using it
    e->cur_tse = e->h_tse[0];
    for (int ci = 0; ci < e->C; ci++) e->cur_chi2[ci] = e->h_chi2[ci];
    e->last_k = 0;
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
    if (int rc0 = flush_commit(e)) return rc0;
    int rc = stage_and_eval(e, k, group_ids, new_box);
    if (rc) return rc;
    memcpy(tse_out, e->h_tse.data(), sizeof(float) * k);
    if (chi2_out) memcpy(chi2_out, e->h_chi2.data(), sizeof(float) * k * e->C);
    return FRMC_OK;
}

int frmc_engine_eval_atoms(frmc_engine *e, int32_t k, const int32_t *counts, const int32_t *atoms,
                           const float *new_box, float *tse_out, float *chi2_out) {
    FRMC_REQUIRE(e, "NULL engine");
    FRMC_REQUIRE(k >= 1 && k <= e->kmax, "k=%d out of range [1,%d]", k, e->kmax);
    FRMC_REQUIRE(counts && atoms && new_box && tse_out, "NULL argument");
    if (int rc0 = flush_commit(e)) return rc0;
    if (int rc0 = update_cells(e)) return rc0;
    int s = 0;
    if (int rc0 = stage_host_atoms(e, k, counts, atoms, new_box, s)) return rc0;
    Stage &S = e->stage[s];
    FRMC_CHECK_HIP(hipMemcpyAsync(S.dev, S.host, S.bytes, hipMemcpyHostToDevice, e->st));
    Inline &I = e->last_inline;
    I.k = 0;
    e->win_dn = e->dnbuf[e->cur_dn];
    e->deferred = false;
    e->device_decided = false;
    int rc = launch_window(e, s, k, &I);
    if (rc) return rc;
    rc = complete_eval(e, k);
    if (rc) return rc;
    memcpy(tse_out, e->h_tse.data(), sizeof(float) * k);
    if (chi2_out) memcpy(chi2_out, e->h_chi2.data(), sizeof(float) * k * e->C);
    return FRMC_OK;
}

int frmc_engine_commit(frmc_engine *e, int32_t p) {
    FRMC_REQUIRE(e, "NULL engine");
    FRMC_REQUIRE(p >= 0 && p < e->last_k, "proposal %d not in the last evaluated window (k=%d)", p, e->last_k);
    FRMC_REQUIRE(!e->device_decided, "the device already committed the last window");
    View V = make_view(e, e->cur_stage);
    CommitCtl C;
    C.k = e->last_k;
    C.p = p;
    C.cur = e->cur_tse;
    C.tol = 0.0f;
    C.dn = e->dnbuf[1 - e->cur_dn];  // the window's deltas (the other buffer is current)
    int rc;
    {
        KTimer t(e, FRMC_K_COMMIT);
        rc = launch_commit(V, e->last_inline, C, e->st);
    }
    if (rc) return rc;
    note_commit(e, p);
    return FRMC_OK;
}

int frmc_engine_sync(frmc_engine *e) {
    FRMC_REQUIRE(e, "NULL engine");
    return finish_windows(e);
}

int frmc_engine_get_state(frmc_engine *e, float *box_out, int32_t constraint, int32_t *slot_out, int32_t *base_out,
                          float *chi2_out, float *tse_out) {
    FRMC_REQUIRE(e, "NULL engine");
    if (int rc0 = flush_commit(e)) return rc0;
    hipStream_t st = e->st;
    if (box_out)
        FRMC_CHECK_HIP(hipMemcpyAsync(box_out, e->xyz, sizeof(float) * 3 * (size_t)e->n, hipMemcpyDeviceToHost, st));
    if (slot_out || base_out) {
        FRMC_REQUIRE(constraint >= 0 && constraint < e->C, "constraint %d out of range", constraint);
        DevC &dc = e->c[constraint];
        if (slot_out)
            FRMC_CHECK_HIP(hipMemcpyAsync(slot_out, dc.slot,
                                          sizeof(int32_t) * (2 * (size_t)e->nel * e->nel * dc.H + 1),
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

int frmc_engine_set_counts(frmc_engine *e, int32_t constraint, const int32_t *slot_in) {
    FRMC_REQUIRE(e && slot_in, "NULL argument");
    if (int rc0 = flush_commit(e)) return rc0;
    FRMC_REQUIRE(constraint >= 0 && constraint < e->C, "constraint %d out of range", constraint);
    DevC &dc = e->c[constraint];
    const size_t count = 2 * (size_t)e->nel * e->nel * dc.H + 1;
    FRMC_CHECK_HIP(hipMemcpyAsync(dc.slot, slot_in, sizeof(int32_t) * count, hipMemcpyHostToDevice, e->st));
    int rc = launch_symmetrise(dc.base, dc.slot, e->nel, dc.H, e->st);
    if (rc) return rc;
    FRMC_CHECK_HIP(hipStreamSynchronize(e->st));
    return FRMC_OK;
}

int frmc_engine_set_real(frmc_engine *e, const float *real_in) {
    FRMC_REQUIRE(e && real_in, "NULL argument");
    e->h_real.assign(real_in, real_in + 3 * (size_t)e->n);
    e->have_real = true;
    return FRMC_OK;
}

int frmc_engine_get_real(frmc_engine *e, float *real_out) {
    FRMC_REQUIRE(e && real_out, "NULL argument");
    FRMC_REQUIRE(e->have_real, "engine was created without real coordinates");
    memcpy(real_out, e->h_real.data(), sizeof(float) * 3 * (size_t)e->n);
    return FRMC_OK;
}

int frmc_engine_set_adaptive(frmc_engine *e, int enable) {
    FRMC_REQUIRE(e, "NULL engine");
    e->adaptive = enable != 0;
    return FRMC_OK;
}

int frmc_engine_profile(frmc_engine *e, int every) {
    FRMC_REQUIRE(e, "NULL engine");
    FRMC_REQUIRE(every >= 0, "every must be >= 0");
    FRMC_CHECK_HIP(hipStreamSynchronize(e->st));
    collect_times(e);
    e->profile = every != 0;
    e->profile_every = every > 0 ? every : 1;
    for (int q = 0; q < FRMC_NKERNELS; q++) {
        e->kernel_ms[q] = 0.0;
        e->kernel_launches[q] = 0;
    }
    return FRMC_OK;
}

int frmc_engine_time_eval(frmc_engine *e, int32_t k, int32_t reps, float *ms_out) {
    FRMC_REQUIRE(e && ms_out, "NULL argument");
    FRMC_REQUIRE(k >= 1 && k <= e->kmax, "k=%d out of range [1,%d]", k, e->kmax);
    FRMC_REQUIRE(reps >= 1 && reps <= 100000, "reps=%d out of range [1,100000]", reps);
    // k_eval_a writes only the window's S(Q) and chi-square scratch, which
    // the next window recomputes: the engine state is untouched
    const int32_t shape = eval_shape(e, k);
    set_units(e, (shape & 16) != 0);
    View V = make_view(e, e->cur_stage);
    V.seq = 0;  // results under no window's number
    V.zero_len = 0;
    V.leafmode = (shape & 16) ? 1 : 0;
    hipEvent_t a = nullptr, b = nullptr;
    FRMC_CHECK_HIP(hipEventCreate(&a));
    hipError_t err = hipEventCreate(&b);
    int rc = FRMC_OK;
    if (err == hipSuccess) err = hipEventRecord(a, e->st);
    for (int32_t r = 0; r < reps && err == hipSuccess && rc == FRMC_OK; r++) rc = launch_eval(V, k, e->st, shape);
    if (err == hipSuccess && rc == FRMC_OK) err = hipEventRecord(b, e->st);
    if (err == hipSuccess && rc == FRMC_OK) err = hipEventSynchronize(b);
    float ms = 0.f;
    if (err == hipSuccess && rc == FRMC_OK) err = hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (rc) return rc;
    if (err != hipSuccess) {
        set_error("k_eval_a timing failed: %s", hipGetErrorString(err));
        return FRMC_EHIP;
    }
    *ms_out = ms / reps;
    return FRMC_OK;
}

int frmc_set_eval_shape(int32_t shape) {
    FRMC_REQUIRE(shape == -1 || (shape >= 0 && shape <= 60 && shape % 4 == 0) || shape == 80 || shape == 84 ||
                     shape == 88 || shape == 92,
                 "eval shape must be -1 (tuned), one of 0, 4, 8, ..., 60, or 80, 84, 88, 92 (got %d)", shape);
    g_eval_forced = shape;
    return FRMC_OK;
}

int frmc_set_delta_cells(int32_t mode) {
    FRMC_REQUIRE(mode >= -1 && mode <= 1, "delta cells mode must be -1 (by size), 0 (off) or 1 (on) (got %d)", mode);
    g_cells_mode = mode;
    return FRMC_OK;
}

int frmc_engine_delta_cells(frmc_engine *e, int32_t *nchunk_out, float *wcut_out) {
    FRMC_REQUIRE(e && nchunk_out, "NULL argument");
    *nchunk_out = e->cells ? e->nchunk : 0;
    if (wcut_out) *wcut_out = e->cells ? e->wcut : 0.0f;
    return FRMC_OK;
}

int frmc_set_window_loop(int32_t mode) {
    FRMC_REQUIRE(mode >= -1 && mode <= 1, "window loop mode must be -1 (default), 0 (off) or 1 (on) (got %d)", mode);
    g_loop_mode = mode;
    return FRMC_OK;
}

int frmc_engine_loop_windows(frmc_engine *e, int64_t *windows_out) {
    FRMC_REQUIRE(e && windows_out, "NULL argument");
    *windows_out = e->loop_windows;
    return FRMC_OK;
}

int frmc_engine_eval_shapes(frmc_engine *e, int32_t *shapes_out, int32_t n) {
    FRMC_REQUIRE(e && shapes_out && n >= 1 && n <= 16, "bad arguments");
    (void)eval_shape(e, 1);  // tunes on first use
    for (int32_t b = 0; b < n; b++)
        shapes_out[b] = g_eval_forced >= 0 ? g_eval_forced : e->eval_shape[b];
    return FRMC_OK;
}

int frmc_engine_kernel_times(frmc_engine *e, double *ms_out, int64_t *launches_out) {
    FRMC_REQUIRE(e && ms_out && launches_out, "NULL argument");
    FRMC_CHECK_HIP(hipStreamSynchronize(e->st));
    collect_times(e);
    for (int q = 0; q < FRMC_NKERNELS; q++) {
        ms_out[q] = e->kernel_ms[q];
        launches_out[q] = e->kernel_launches[q];
    }
    return FRMC_OK;
}

}  // extern "C"
