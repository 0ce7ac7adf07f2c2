// driver.hip -- native run loop for single-group translation moves.
//
// Reproduces fullrmc's proposal stream exactly from its two MT19937
// generators (CPython's random: group index, amplitude, acceptance draw;
// NumPy's legacy RandomState: direction), evaluates k proposals per window
// on the device, walks them in order with the reference's accept rule, and
// commits the first accepted one after rewinding both streams to just after
// its draws.  Every proposal before the first acceptance was rejected with
// exactly one acceptance draw (Engine.py:3308-3317 draws only when the total
// got worse, and with the default tolerance a worse move is always
// rejected), so the k proposals of a window are the ones the sequential loop
// generates, and the trajectory is the sequential one.
#include <math.h>
#include <sched.h>
#include <time.h>

#include <algorithm>
#include <atomic>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "draws.h"
#include "engine.h"

namespace frmc {

// MT19937 (Matsumoto & Nishimura 1998); `used` counts 32-bit words drawn so a
// stream can be rewound to any proposal boundary by replaying that many.
struct Stream {
    frmc_mt19937 *s;
    int32_t used;
    inline uint32_t next() {
        if (s->pos >= kMtN) {
            for (int i = 0; i < kMtN; i++) s->mt[i] = s->mt[(i + 397) % kMtN] ^ mt_mix(s->mt[i], s->mt[(i + 1) % kMtN]);
            s->pos = 0;
        }
        used++;
        return mt_temper(s->mt[s->pos++]);
    }
    // random.random() and RandomState.random_sample(): 53-bit double
    inline double uniform() {
        const uint32_t a = next(), b = next();
        return mt_uniform(a, b);
    }
    // random.randint(0, n-1) == randrange(0, n) -> _randbelow_with_getrandbits(n)
    inline int32_t randbelow(int32_t n) {
        const int k = bit_length(n);
        uint32_t r;
        do {
            r = next() >> (32 - k);
        } while (r >= (uint32_t)n);
        return (int32_t)r;
    }
};

// generate_random_vector(minAmp, maxAmp) (Core/Collection.py:674-701)
static inline void random_vector(Stream &py, Stream &np_, float minAmp, float maxAmp, float v[3]) {
    float norm;
    double u[3];
    do {
        for (int d = 0; d < 3; d++) u[d] = np_.uniform();
    } while (!random_direction(u, v, norm));
    scale_direction(v, norm, py.uniform(), minAmp, maxAmp);
}

constexpr int32_t kReplicaThreads = 16;  // host threads driving independent replicas

// Window size from the running acceptance rate: enough proposals that the
// window holds an acceptance with ~90 % probability, within [kmin, kmax].
// (90 % measured 2-3 % faster than 95 % on the 13,500-atom workload, and
// 98 % slower: fewer proposals evaluated after the first acceptance against
// more windows; the trajectory does not depend on it.)
static int32_t window_size(int64_t generated, int64_t accepted, int32_t kmax) {
    const double p = (accepted + 1.0) / (generated + 2.0);
    double k = p >= 1.0 ? 1.0 : log(0.10) / log(1.0 - p);
    const int32_t kmin = std::min<int32_t>(kmax, 4);
    int32_t kk = (int32_t)std::ceil(k);
    return std::max(kmin, std::min(kmax, kk));
}


// One engine's native translation run, advanced one window at a time:
// launch() draws the window's proposals and launches their evaluation (and
// the device-decided commit), finish() waits for the totals, walks them in
// order with the reference's accept rule and rewinds both streams to just
// after the committed proposal's draws.  Several runs can have windows in
// flight at once (frmc_engines_run_translation).
struct TranslationRun {
    frmc_engine *e = nullptr;
    int64_t nsteps = 0;
    int32_t k = 0;
    float minAmp = 0.f, maxAmp = 0.f, tolerance = 0.f;
    frmc_mt19937 *py_state = nullptr, *np_state = nullptr;
    frmc_run_stats *stats = nullptr;
    int32_t *group_log = nullptr;
    uint8_t *accept_log = nullptr;
    float *tse_log = nullptr;
    std::vector<int32_t> gids, aoff, py_pre, py_post, np_used;
    std::vector<float> nbox, nreal, uf, tse;
    std::vector<double> uacc;
    int64_t step = 0, gen0 = 0, acc0 = 0;
    int32_t kw = 0;
    frmc_mt19937 py0, np0;

    int init() {
        FRMC_REQUIRE(e && py_state && np_state && stats, "NULL argument");
        FRMC_REQUIRE(e->have_real, "engine was created without real coordinates");
        FRMC_REQUIRE(k >= 1 && k <= e->kmax, "k=%d out of range [1,%d]", k, e->kmax);
        FRMC_REQUIRE(py_state->pos >= 0 && py_state->pos <= 624 && np_state->pos >= 0 && np_state->pos <= 624,
                     "bad MT19937 position");
        gids.resize(k);
        aoff.resize(k + 1);
        py_pre.resize(k);
        py_post.resize(k);
        np_used.resize(k);
        nbox.resize((size_t)k * e->gmax * 3);
        nreal.resize((size_t)k * e->gmax * 3);
        uf.resize(k);
        tse.resize(k);
        uacc.resize(k);
        return FRMC_OK;
    }

    bool done() const { return step >= nsteps; }

    // proposals in the next window
    int32_t next_k() const {
        const int32_t kk = e->adaptive ? window_size(gen0, acc0, k) : k;
        return (int32_t)std::min<int64_t>(kk, nsteps - step);
    }

    int launch() {
        kw = next_k();
        draw();
        return launch_device_decided(e, kw, gids.data(), nbox.data(), uf.data(), tolerance);
    }

    // draw the kw proposals of the next window from both streams (saved first
    // so that finish() can rewind them)
    void draw() {
        py0 = *py_state;
        np0 = *np_state;
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
            uacc[p] = py.uniform();  // the acceptance draw, consumed only if the move is worse
            py_post[p] = py.used;
            aoff[p] = off;
            const int32_t g0 = e->h_goff[gi], g = e->h_goff[gi + 1] - g0;
            for (int32_t a = 0; a < g; a++) {
                const size_t id = (size_t)e->h_gatoms[g0 + a];
                float r[3], b[3];
                for (int d = 0; d < 3; d++) r[d] = e->h_real[3 * id + d] + v[d];
                to_box(e->recip, r, b);
                for (int d = 0; d < 3; d++) {
                    nreal[3 * (size_t)(off + a) + d] = r[d];
                    nbox[3 * (size_t)(off + a) + d] = b[d];
                }
            }
            off += g;
        }
        aoff[kw] = off;
        for (int32_t p = 0; p < kw; p++) uf[p] = (float)uacc[p];
    }

    // Batches of windows on the device loop while a whole window is left:
    // k is fixed per batch (from the acceptance rate so far), and batches
    // start small so that it adapts early.  The host loop runs what is left.
    int run_loop() {
        if (!loop_usable(e, k)) return FRMC_OK;
        if (int rc = loop_begin(e, *py_state, *np_state)) return rc;
        int32_t ramp = 16;
        while (!done()) {
            const int32_t kk = e->adaptive ? window_size(gen0, acc0, k) : k;
            const int64_t fit = (nsteps - step) / kk;
            int32_t M = 0;
            if (int rc = loop_plan(e, maxAmp, (int32_t)std::min<int64_t>(std::min(ramp, kLoopWindows), fit), M))
                return rc;
            if (M < 1) break;
            if (int rc = loop_batch(e, kk, M, minAmp, maxAmp, tolerance)) return rc;
            const LoopState &L = *e->h_loop;
            const int64_t walked = L.h.step;
            stats->windows += M;
            stats->evaluated += (int64_t)M * kk;
            stats->generated += walked;
            stats->tried += walked;
            stats->accepted += L.h.nacc;
            stats->tolerated += L.h.ntol;
            gen0 += walked;
            acc0 += L.h.nacc;
            for (int64_t i = 0; i < walked; i++) {
                if (group_log) group_log[step + i] = e->h_loop_glog[i];
                if (accept_log) accept_log[step + i] = e->h_loop_alog[i];
                if (tse_log) tse_log[step + i] = e->h_loop_tlog[i];
            }
            step += walked;
            memcpy(py_state->mt, L.py.mt, sizeof(py_state->mt));
            py_state->pos = L.py.pos;
            memcpy(np_state->mt, L.np.mt, sizeof(np_state->mt));
            np_state->pos = L.np.pos;
            ramp = std::min(2 * ramp, kLoopWindows);
        }
        return FRMC_OK;
    }

    int finish() {
        int rc = finish_device_decided(e, kw, tse.data());
        if (rc) return rc;
        stats->windows++;
        stats->evaluated += kw;
        // walk the window in order (Engine.py:3308-3338)
        int32_t acc_p = -1;
        bool drew = false;
        for (int32_t p = 0; p < kw; p++) {
            stats->generated++;
            stats->tried++;
            gen0++;
            bool accept = true;
            drew = false;
            if (tse[p] > e->cur_tse) {
                drew = true;
                // random() > tolerance, the Python float compared as float32
                if ((float)uacc[p] > tolerance) accept = false;
                else stats->tolerated++;
            }
            if (group_log) group_log[step] = gids[p];
            if (accept_log) accept_log[step] = accept ? 1 : 0;
            if (tse_log) tse_log[step] = accept ? tse[p] : e->cur_tse;
            step++;
            if (accept) {
                stats->accepted++;
                acc0++;
                acc_p = p;
                break;
            }
        }
        if (acc_p >= 0) {
            // rewind both streams to just after the accepted proposal's draws
            *py_state = py0;
            *np_state = np0;
            Stream rp{py_state, 0}, rn{np_state, 0};
            const int32_t npy = drew ? py_post[acc_p] : py_pre[acc_p];
            while (rp.used < npy) (void)rp.next();
            while (rn.used < np_used[acc_p]) (void)rn.next();
            rc = settle_window(e, acc_p);
            if (rc) return rc;
            const int32_t gi = gids[acc_p], g0 = e->h_goff[gi], g = e->h_goff[gi + 1] - g0;
            for (int32_t a = 0; a < g; a++) {
                const size_t id = (size_t)e->h_gatoms[g0 + a];
                for (int d = 0; d < 3; d++) e->h_real[3 * id + d] = nreal[3 * (size_t)(aoff[acc_p] + a) + d];
            }
            return FRMC_OK;
        }
        // without an acceptance both streams already stand after the window
        return settle_window(e, -1);
    }
};

static int g_batch_replicas = 1;  // frmc_set_replica_batching
static int g_replica_threads = 0;  // frmc_set_replica_threads (0: automatic)

// host threads for replicas: the CPUs this process may run on, less two for
// the caller's other threads and the HIP runtime, at most kReplicaThreads
static int32_t replica_threads(int32_t nengines) {
    int32_t cpus = (int32_t)std::max(1u, std::thread::hardware_concurrency());
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) cpus = std::max(1, CPU_COUNT(&set));
    int32_t T = g_replica_threads > 0 ? g_replica_threads : std::min<int32_t>(kReplicaThreads, std::max(1, cpus - 2));
    return std::max<int32_t>(1, std::min<int32_t>(T, nengines));
}
// host time of the last batched run, by phase (frmc_replica_batch_times):
// drawing and staging, upload and launch, waiting and walking, set-up,
// final commits; steps
static double g_batch_ms[5] = {0, 0, 0, 0, 0};
static int64_t g_batch_steps = 0;

static double now_ms() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + 1e-6 * ts.tv_nsec;
}

// Fixed worker threads for the host halves of a batched step: run(f) calls
// f(t) on every thread t in [0, T) (t = 0 on the caller) and returns when all
// have finished.  Workers spin between steps (a step is tens of microseconds).
class StepPool {
  public:
    explicit StepPool(int32_t T) : T_(T) {
        for (int32_t t = 1; t < T_; t++) th_.emplace_back([this, t] { loop(t); });
    }
    ~StepPool() {
        stop_.store(true);
        gen_.fetch_add(1);
        for (auto &x : th_) x.join();
    }
    void run(const std::function<void(int32_t)> &f) {
        job_ = &f;
        left_.store(T_ - 1);
        gen_.fetch_add(1);
        f(0);
        while (left_.load() > 0) std::this_thread::yield();
    }

  private:
    void loop(int32_t t) {
        uint64_t seen = 0;
        for (;;) {
            uint64_t g;
            while ((g = gen_.load()) == seen) std::this_thread::yield();
            seen = g;
            if (stop_.load()) return;
            (*job_)(t);
            left_.fetch_sub(1);
        }
    }
    int32_t T_;
    std::vector<std::thread> th_;
    std::atomic<uint64_t> gen_{0};
    std::atomic<int32_t> left_{0};
    std::atomic<bool> stop_{false};
    const std::function<void(int32_t)> *job_ = nullptr;
};

static size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

constexpr int32_t kLanes = 2;  // replica groups whose steps alternate on the GPU

// One lane of a batched run: a fixed set of replicas and its own host and
// device step blocks (windows, block maps, proposals), so that one lane's
// host work overlaps the other lane's kernels.
struct Lane {
    std::vector<int32_t> reps;       // replicas of the lane
    std::vector<int32_t> act, kws;  // replica and window size of each step slot
    int32_t nact = 0;
    bool inflight = false;
    size_t win_b = 0, map_b = 0, bytes = 0;
    void *h = nullptr, *d = nullptr;
    BatchArgs A{};
    int32_t nd = 0, na = 0, nb = 0;
};

// All replicas' windows in one launch per kernel per step and lane.  In one
// parallel pass over a lane's replicas the host walks each replica's totals
// of the window in flight with the reference's rule, exactly as the
// single-engine loop does, and draws and stages its next window; then it
// uploads windows, proposals and block maps in one copy and launches
// k_delta_batch, k_eval_a_batch and (when some window needs it)
// k_eval_b_batch on the first engine's stream.  The replicas are split into
// two lanes whose steps alternate, so one lane's host work runs while the
// other's kernels do.  Each replica keeps its own state, streams and
// deferred commit, so every trajectory is the one its engine alone would
// produce.
static int run_batched(std::vector<TranslationRun> &runs, int32_t T) {
    const double t_start = now_ms();
    const int32_t R = (int32_t)runs.size();
    hipStream_t st = runs[0].e->st;
    // the engines' own streams may still hold work on their buffers
    for (auto &r : runs) {
        if (int rc = flush_commit(r.e)) return rc;
        FRMC_CHECK_HIP(hipStreamSynchronize(r.e->st));
    }
    std::vector<View> hv(R);
    size_t shm_a = 0, shm_b = 0, shm_d = 0;
    int32_t kmax = 1;
    int64_t maxd = 0, maxa = 0;
    for (int32_t r = 0; r < R; r++) {
        use_unit_layout(runs[r].e, false);  // the batched kernels take 64-Q units and k_eval_b_batch
        drop_cells(runs[r].e);              // k_delta_batch walks all atoms and does not keep the sorted copy
        hv[r] = make_view(runs[r].e, 0);
        if (kBP * hv[r].dnStride <= kBatchLdsInts)
            shm_d = std::max(shm_d, sizeof(int32_t) * kBP * hv[r].dnStride);
        shm_a = std::max(shm_a, eval_lds_bytes(hv[r]));
        shm_b = std::max(shm_b, eval_b_lds_bytes(hv[r]));
        kmax = std::max(kmax, runs[r].k);
        maxd += (int64_t)((hv[r].n + kChunk - 1) / kChunk) * ((runs[r].k + kBP - 1) / kBP);
        maxa += (int64_t)hv[r].nunits * runs[r].k;
    }
    FRMC_REQUIRE(maxd < ((int64_t)1 << 30) && maxa < ((int64_t)1 << 30), "batched grid too large");
    std::vector<Lane> lanes(std::min<int32_t>(kLanes, R));
    const int32_t L = (int32_t)lanes.size();
    // replicas dealt to the lanes in order of size, so the lanes carry equal work
    {
        std::vector<int32_t> order(R);
        for (int32_t r = 0; r < R; r++) order[r] = r;
        std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return hv[x].n > hv[y].n; });
        for (int32_t q = 0; q < R; q++) lanes[q % L].reps.push_back(order[q]);
    }
    View *d_views = nullptr;
    struct Free {
        hipStream_t st;
        View *&v;
        std::vector<Lane> &ls;
        ~Free() {
            (void)hipStreamSynchronize(st);  // nothing in flight reads the blocks
            if (v) (void)hipFree(v);
            for (auto &l : ls) {
                if (l.d) (void)hipFree(l.d);
                if (l.h) (void)hipHostFree(l.h);
            }
        }
    } guard{st, d_views, lanes};
    FRMC_CHECK_HIP(hipMalloc((void **)&d_views, sizeof(View) * R));
    FRMC_CHECK_HIP(hipMemcpyAsync(d_views, hv.data(), sizeof(View) * R, hipMemcpyHostToDevice, st));
    // step block of a lane: windows, then the step's proposals, then the three
    // block maps
    for (auto &l : lanes) {
        const int32_t n = (int32_t)l.reps.size();
        int64_t md = 0, ma = 0;
        for (int32_t r : l.reps) {
            md += (int64_t)((hv[r].n + kChunk - 1) / kChunk) * ((runs[r].k + kBP - 1) / kBP);
            ma += (int64_t)hv[r].nunits * runs[r].k;
        }
        l.win_b = align16(sizeof(BatchWin) * n);
        l.map_b = align16(sizeof(int2) * (md + 2 * ma));
        l.bytes = l.win_b + align16(sizeof(Prop) * (size_t)n * kmax) + l.map_b;
        l.act.resize(n);
        l.kws.resize(n);
        FRMC_CHECK_HIP(hipMalloc(&l.d, l.bytes));
        FRMC_CHECK_HIP(hipHostMalloc(&l.h, l.bytes, hipHostMallocDefault));
        l.A.views = d_views;
        l.A.win = (const BatchWin *)l.d;
    }
    std::vector<int> rcs(T, FRMC_OK);
    std::vector<std::string> errs(T);
    auto fail = [&](int32_t t, int rc) {
        if (rc && !rcs[t]) {
            rcs[t] = rc;
            errs[t] = frmc_last_error();
        }
    };
    auto first_error = [&]() {
        for (int32_t t = 0; t < T; t++)
            if (rcs[t]) {
                set_error("%s", errs[t].c_str());
                return rcs[t];
            }
        return (int)FRMC_OK;
    };
    // One parallel pass per lane step: each thread takes its replicas of the
    // lane in turn, walks the totals of the window in flight (if any), then
    // draws and stages the replica's next window at the next free slot of
    // the step's window and proposal tables (atomic counters).
    Lane *cur = nullptr;
    std::atomic<int32_t> nslot{0}, noff{0};
    const std::function<void(int32_t)> step = [&](int32_t t) {
        Lane &l = *cur;
        BatchWin *hw = (BatchWin *)l.h;
        Prop *hp = (Prop *)((char *)l.h + l.win_b);
        const int32_t nr = (int32_t)l.reps.size();
        for (int32_t i = t; i < nr && !rcs[t]; i += T) {
            const int32_t r = l.reps[i];
            TranslationRun &run = runs[r];
            if (run.kw > 0) {
                fail(t, run.finish());
                run.kw = 0;
                if (rcs[t]) break;
            }
            if (run.done()) continue;
            const int32_t kw = run.next_k();
            const int32_t slot = nslot.fetch_add(1), off = noff.fetch_add(kw);
            run.kw = kw;
            run.draw();
            BatchWin &W = hw[slot];
            W.rep = r;
            W.prop = off;
            l.act[slot] = r;
            l.kws[slot] = kw;
            fail(t, batch_prepare(run.e, kw, run.gids.data(), run.nbox.data(), W, hp + off));
        }
    };
    StepPool pool(T);
    for (double &x : g_batch_ms) x = 0.0;
    g_batch_steps = 0;
    // walk lane l's window in flight, draw and stage the next one, upload and
    // launch it
    auto step_lane = [&](Lane &l) -> int {
        const double t0 = now_ms();
        cur = &l;
        nslot.store(0);
        noff.store(0);
        pool.run(step);
        if (int rc = first_error()) return rc;
        l.nact = nslot.load();
        const int32_t off = noff.load();
        l.inflight = l.nact > 0;
        if (!l.inflight) return FRMC_OK;
        g_batch_steps++;
        const double t1 = now_ms();
        // block maps: (window, block within the window) per block of each
        // kernel, right after the step's proposals
        const size_t moff = l.win_b + align16(sizeof(Prop) * (size_t)off);
        int2 *hm = (int2 *)((char *)l.h + moff);
        int32_t nd = 0, na = 0, nb = 0;
        for (int32_t i = 0; i < l.nact; i++) {
            const View &V = hv[l.act[i]];
            const int32_t k = l.kws[i], bd = ((V.n + kChunk - 1) / kChunk) * ((k + kBP - 1) / kBP);
            for (int32_t q = 0; q < bd; q++) hm[nd++] = make_int2(i, q);
        }
        for (int32_t i = 0; i < l.nact; i++) {
            const View &V = hv[l.act[i]];
            const int32_t ba = V.nunits * l.kws[i];
            for (int32_t q = 0; q < ba; q++) hm[nd + na++] = make_int2(i, q);
        }
        for (int32_t i = 0; i < l.nact; i++) {
            const View &V = hv[l.act[i]];
            if (V.C == 1 && V.c[0].rspace) continue;  // k_eval_a_batch publishes the total
            for (int32_t q = 0; q < l.kws[i]; q++) hm[nd + na + nb++] = make_int2(i, q);
        }
        l.A.props = (const Prop *)((char *)l.d + l.win_b);
        l.A.map_d = (const int2 *)((char *)l.d + moff);
        l.A.map_a = l.A.map_d + nd;
        l.A.map_b = l.A.map_a + na;
        // windows, proposals and maps: one copy of the used bytes
        FRMC_CHECK_HIP(hipMemcpyAsync(l.d, l.h, moff + sizeof(int2) * (nd + na + nb), hipMemcpyHostToDevice, st));
        if (int rc = launch_batch(l.A, nd, na, nb, shm_d, shm_a, shm_b, st)) return rc;
        g_batch_ms[0] += t1 - t0;
        g_batch_ms[1] += now_ms() - t1;
        return FRMC_OK;
    };
    g_batch_ms[3] = now_ms() - t_start;
    for (auto &l : lanes) l.inflight = true;  // the first pass only draws
    for (;;) {
        bool any = false;
        for (auto &l : lanes) {
            if (!l.inflight) continue;
            any = true;
            if (int rc = step_lane(l)) return rc;
        }
        if (!any) break;
    }
    // the accepted moves of the last windows, on the batch stream
    const double t_end = now_ms();
    for (auto &r : runs)
        if (int rc = flush_commit(r.e, st)) return rc;
    FRMC_CHECK_HIP(hipStreamSynchronize(st));
    for (auto &r : runs)
        if (int rc = finish_windows(r.e)) return rc;
    g_batch_ms[4] = now_ms() - t_end;
    return FRMC_OK;
}

// Batched replicas on the device window loop: per step one
// k_loop_step_batch (every replica's walk of its previous window and draw of
// its next, one workgroup each) and the batched delta and evaluation
// kernels, queued for a batch of M steps with window sizes fixed per replica
// for the batch; the host uploads the batch's tables once and synchronises
// once per batch, so it is off the per-step critical path.  A replica's last
// windows draw only the moves it has left (as the host loop's last window
// does), and a replica that has finished walks empty windows until the
// batch ends, so the batches run every replica to its end.  The replicas'
// real coordinates and result words live in the run's own device block.
static int run_batched_loop(std::vector<TranslationRun> &runs) {
    const int32_t R = (int32_t)runs.size();
    for (auto &r : runs)
        if (!loop_batch_usable(r.e, r.k)) return FRMC_OK;
    hipStream_t st = runs[0].e->st;
    // the engines' own streams may still hold work on their buffers; a
    // pending commit goes to the loop state (the first step's k_delta_batch
    // applies it)
    for (auto &r : runs) FRMC_CHECK_HIP(hipStreamSynchronize(r.e->st));
    std::vector<View> hv(R);
    std::vector<size_t> o_real(R), o_res(R);
    size_t shm_a = 0, shm_b = 0, shm_d = 0, shm_l = 4, n_real = 0, n_res = 0;
    int32_t kmax = 1;
    int64_t maxd = 0, maxa = 0, maxb = 0;
    for (int32_t r = 0; r < R; r++) {
        frmc_engine *e = runs[r].e;
        use_unit_layout(e, false);  // 64-Q units, totals from k_eval_b_batch (or k_eval_a_batch)
        drop_cells(e);
        if (e->groups_ident < 0) {
            bool id = e->G == e->n;
            for (int32_t g = 0; g < e->G && id; g++) id = e->h_goff[g] == g && e->h_gatoms[g] == g;
            e->groups_ident = id ? 1 : 0;
        }
        hv[r] = make_view(e, 0);
        o_real[r] = n_real;
        n_real += 3 * (size_t)e->n;
        o_res[r] = n_res;
        n_res += (size_t)runs[r].k * e->resW;
        if (kBP * hv[r].dnStride <= kBatchLdsInts) shm_d = std::max(shm_d, sizeof(int32_t) * kBP * hv[r].dnStride);
        shm_a = std::max(shm_a, eval_lds_bytes(hv[r]));
        shm_b = std::max(shm_b, eval_b_lds_bytes(hv[r]));
        shm_l = std::max(shm_l, sizeof(float) * (size_t)runs[r].k * e->resW);
        kmax = std::max(kmax, runs[r].k);
        maxd += (int64_t)((hv[r].n + kChunk - 1) / kChunk) * ((runs[r].k + kBP - 1) / kBP);
        maxa += (int64_t)hv[r].nunits * runs[r].k;
        maxb += runs[r].k;
    }
    FRMC_REQUIRE(maxd + maxa + maxb < ((int64_t)1 << 30), "batched grid too large");
    // one device block and its pinned mirror: loop states, arguments,
    // accepted moves, proposals, the three window tables, the block maps,
    // the views, the real coordinates, the result words
    const size_t b_states = align16(sizeof(LoopState) * R), b_args = align16(sizeof(LoopArgs) * R),
                 b_acc = align16(sizeof(LoopAcc) * (size_t)R * kLoopWindows),
                 b_props = align16(sizeof(Prop) * ((size_t)R * kmax + kLoopMaxK)),
                 b_wins = align16(sizeof(BatchWin) * 3 * (size_t)R),
                 b_maps = align16(sizeof(int2) * (size_t)(maxd + maxa + maxb)), b_views = align16(sizeof(View) * R),
                 b_real = align16(sizeof(float) * n_real), b_res = align16(sizeof(unsigned long long) * n_res);
    const size_t o_args = b_states, o_acc = o_args + b_args, o_props = o_acc + b_acc, o_wins = o_props + b_props,
                 o_maps = o_wins + b_wins, o_views = o_maps + b_maps, o_reals = o_views + b_views,
                 o_ress = o_reals + b_real, bytes = o_ress + b_res;
    char *d = nullptr, *h = nullptr;
    struct Free {
        hipStream_t st;
        char *&d, *&h;
        ~Free() {
            (void)hipStreamSynchronize(st);
            if (d) (void)hipFree(d);
            if (h) (void)hipHostFree(h);
        }
    } guard{st, d, h};
    FRMC_CHECK_HIP(hipMalloc((void **)&d, bytes));
    FRMC_CHECK_HIP(hipHostMalloc((void **)&h, bytes, hipHostMallocDefault));
    LoopState *d_states = (LoopState *)d, *h_states = (LoopState *)h;
    LoopArgs *d_args = (LoopArgs *)(d + o_args), *h_args = (LoopArgs *)(h + o_args);
    LoopAcc *d_acc = (LoopAcc *)(d + o_acc), *h_acc = (LoopAcc *)(h + o_acc);
    Prop *d_props = (Prop *)(d + o_props);
    BatchWin *d_wins = (BatchWin *)(d + o_wins), *h_wins = (BatchWin *)(h + o_wins);
    int2 *d_maps = (int2 *)(d + o_maps), *h_maps = (int2 *)(h + o_maps);
    View *d_views = (View *)(d + o_views), *h_views = (View *)(h + o_views);
    float *d_real = (float *)(d + o_reals), *h_real = (float *)(h + o_reals);
    unsigned long long *d_res = (unsigned long long *)(d + o_ress);
    for (int32_t r = 0; r < R; r++) {
        frmc_engine *e = runs[r].e;
        loop_init_state(e, h_states[r], *runs[r].py_state, *runs[r].np_state);
        hv[r].pend_dev = &d_states[r].h.pend;
        hv[r].out_res = d_res + o_res[r];
        h_views[r] = hv[r];
        memcpy(h_real + o_real[r], e->h_real.data(), sizeof(float) * 3 * (size_t)e->n);
    }
    FRMC_CHECK_HIP(hipMemcpyAsync(d, h, b_states, hipMemcpyHostToDevice, st));
    FRMC_CHECK_HIP(hipMemcpyAsync(d_views, h_views, sizeof(View) * R, hipMemcpyHostToDevice, st));
    FRMC_CHECK_HIP(hipMemcpyAsync(d_real, h_real, sizeof(float) * n_real, hipMemcpyHostToDevice, st));
    std::vector<int32_t> kr(R), poff(R), c0(R);
    int32_t ramp = 16;
    for (;;) {
        // the batch: window sizes from each replica's acceptance so far, and
        // about as many steps as the replica furthest from its end needs
        int64_t need = 0;
        int32_t off = 0;
        for (int32_t r = 0; r < R; r++) {
            const TranslationRun &run = runs[r];
            kr[r] = run.e->adaptive ? window_size(run.gen0, run.acc0, run.k) : run.k;
            const double p = (run.acc0 + 1.0) / (run.gen0 + 2.0);
            const int64_t left = run.nsteps - run.step;
            if (left > 0)
                need = std::max<int64_t>(need, std::min<int64_t>(left, (int64_t)std::ceil(left * p * 1.2) + 2));
            poff[r] = off;
            off += kr[r];
            c0[r] = run.e->cur_dn;
        }
        if (need == 0) break;
        const int32_t M = (int32_t)std::min<int64_t>({(int64_t)ramp, (int64_t)kLoopWindows, need});
        // arguments, window tables (first step, odd steps, even steps), block maps
        FRMC_CHECK_HIP(hipStreamSynchronize(st));  // the previous batch no longer reads the pinned block
        int32_t nd = 0, na = 0, nb = 0;
        for (int32_t r = 0; r < R; r++) {
            frmc_engine *e = runs[r].e;
            LoopArgs &A = h_args[r];
            A = LoopArgs{};
            A.L = d_states + r;
            A.resW = e->resW;
            A.rs = e->resW;
            A.C = e->C;
            A.dnStride = e->dnStride;
            A.tol = runs[r].tolerance;
            A.res = d_res + o_res[r];
            A.G = e->G;
            A.nbits = bit_length(e->G);
            A.ident = e->groups_ident;
            A.minAmp = runs[r].minAmp;
            A.maxAmp = runs[r].maxAmp;
            for (int q = 0; q < 9; q++) A.recip[q] = e->recip[q];
            A.goff = e->goff;
            A.gatoms = e->gatoms;
            A.elem = e->elem;
            A.spos = nullptr;
            A.xyz = e->xyz;
            A.real = d_real + o_real[r];
            A.props = d_props + poff[r];
            A.acc = d_acc + (size_t)r * kLoopWindows;
            A.kr = kr[r];
            A.c0 = c0[r];
            A.dnb[0] = e->dnbuf[0];
            A.dnb[1] = e->dnbuf[1];
            A.left = runs[r].nsteps - runs[r].step;
            for (int t = 0; t < 3; t++) {
                BatchWin &W = h_wins[(size_t)t * R + r];
                const int32_t s = t;  // a step of the table's kind: the first, odd, even
                W.rep = r;
                W.k = kr[r];
                W.prop = poff[r];
                W.seq = 0;
                W.dn = e->dnbuf[(c0[r] + s) & 1];
                W.dn_zero = e->dnbuf[(c0[r] + s + 1) & 1];
                W.zero_len = (int64_t)(s == 0 ? e->dn_used[(c0[r] + 1) & 1] : kr[r]) * e->dnStride;
                W.P = Pending{};
                W.P.valid = 0;
            }
        }
        for (int32_t r = 0; r < R; r++) {
            const int32_t bd = ((hv[r].n + kChunk - 1) / kChunk) * ((kr[r] + kBP - 1) / kBP);
            for (int32_t q = 0; q < bd; q++) h_maps[nd++] = make_int2(r, q);
        }
        for (int32_t r = 0; r < R; r++)
            for (int32_t q = 0; q < hv[r].nunits * kr[r]; q++) h_maps[nd + na++] = make_int2(r, q);
        for (int32_t r = 0; r < R; r++) {
            if (hv[r].C == 1 && hv[r].c[0].rspace) continue;  // k_eval_a_batch publishes the total
            for (int32_t q = 0; q < kr[r]; q++) h_maps[nd + na + nb++] = make_int2(r, q);
        }
        FRMC_CHECK_HIP(hipMemcpyAsync(d_args, h_args, sizeof(LoopArgs) * R, hipMemcpyHostToDevice, st));
        FRMC_CHECK_HIP(hipMemcpyAsync(d_wins, h_wins, sizeof(BatchWin) * 3 * R, hipMemcpyHostToDevice, st));
        FRMC_CHECK_HIP(hipMemcpyAsync(d_maps, h_maps, sizeof(int2) * (nd + na + nb), hipMemcpyHostToDevice, st));
        for (int32_t s = 0; s <= M; s++) {
            if (int rc = launch_loop_step_batch(d_args, R, s, M, shm_l, st)) return rc;
            if (s == M) break;
            BatchArgs BA{};
            BA.views = d_views;
            BA.win = d_wins + (size_t)(s == 0 ? 0 : (s & 1) ? 1 : 2) * R;
            BA.props = d_props;
            BA.map_d = d_maps;
            BA.map_a = d_maps + nd;
            BA.map_b = d_maps + nd + na;
            if (int rc = launch_batch(BA, nd, na, nb, shm_d, shm_a, shm_b, st)) return rc;
        }
        // each replica's streams and header, and its accepted moves
        FRMC_CHECK_HIP(hipMemcpy2DAsync(h_states, sizeof(LoopState), d_states, sizeof(LoopState),
                                        offsetof(LoopState, gid), R, hipMemcpyDeviceToHost, st));
        FRMC_CHECK_HIP(hipMemcpy2DAsync(h_acc, sizeof(LoopAcc) * kLoopWindows, d_acc, sizeof(LoopAcc) * kLoopWindows,
                                        sizeof(LoopAcc) * M, R, hipMemcpyDeviceToHost, st));
        FRMC_CHECK_HIP(hipStreamSynchronize(st));
        g_batch_steps += M;
        for (int32_t r = 0; r < R; r++) {
            TranslationRun &run = runs[r];
            frmc_engine *e = run.e;
            const LoopState &L = h_states[r];
            if (int rc = loop_apply(e, L, h_acc + (size_t)r * kLoopWindows, M, (int64_t)M * kr[r])) return rc;
            const int64_t walked = L.h.step;
            FRMC_REQUIRE(walked <= run.nsteps - run.step, "replica %d walked past its end", r);
            run.stats->windows += L.h.nwin;
            run.stats->evaluated += L.h.neval;
            run.stats->generated += walked;
            run.stats->tried += walked;
            run.stats->accepted += L.h.nacc;
            run.stats->tolerated += L.h.ntol;
            run.gen0 += walked;
            run.acc0 += L.h.nacc;
            run.step += walked;
            memcpy(run.py_state->mt, L.py.mt, sizeof(run.py_state->mt));
            run.py_state->pos = L.py.pos;
            memcpy(run.np_state->mt, L.np.mt, sizeof(run.np_state->mt));
            run.np_state->pos = L.np.pos;
            // window s wrote buffer (c0 + s) & 1 and zeroed the other
            e->dn_used[(c0[r] + M - 1) & 1] = kr[r];
            e->dn_used[(c0[r] + M) & 1] = 0;
            e->cur_dn = (int)((c0[r] + M) & 1);
            e->loop_windows += L.h.nwin;
        }
        ramp = std::min(2 * ramp, kLoopWindows);
    }
    return FRMC_OK;
}

}  // namespace frmc

using namespace frmc;

extern "C" void frmc_set_replica_batching(int enable) { g_batch_replicas = enable != 0; }

extern "C" void frmc_set_replica_threads(int nthreads) { g_replica_threads = nthreads > 0 ? nthreads : 0; }

extern "C" int frmc_replica_batch_times(double *ms_out, int64_t *steps_out) {
    FRMC_REQUIRE(ms_out && steps_out, "NULL pointer argument");
    for (int q = 0; q < 5; q++) ms_out[q] = g_batch_ms[q];
    *steps_out = g_batch_steps;
    return FRMC_OK;
}

extern "C" int frmc_engine_run_translation(frmc_engine *e, int64_t nsteps, int32_t k, float minAmp, float maxAmp,
                                           float tolerance, frmc_mt19937 *py_state, frmc_mt19937 *np_state,
                                           frmc_run_stats *stats, int32_t *group_log, uint8_t *accept_log,
                                           float *tse_log) {
    TranslationRun run;
    run.e = e;
    run.nsteps = nsteps;
    run.k = k;
    run.minAmp = minAmp;
    run.maxAmp = maxAmp;
    run.tolerance = tolerance;
    run.py_state = py_state;
    run.np_state = np_state;
    run.stats = stats;
    run.group_log = group_log;
    run.accept_log = accept_log;
    run.tse_log = tse_log;
    int rc = run.init();
    if (rc) return rc;
    if ((rc = run.run_loop())) return rc;
    while (!run.done()) {
        if ((rc = run.launch())) return rc;
        if ((rc = run.finish())) return rc;
    }
    return finish_windows(e);
}

extern "C" int frmc_engines_run_translation(frmc_engine **engines, int32_t nengines, int64_t nsteps, int32_t k,
                                            float minAmp, float maxAmp, float tolerance, frmc_mt19937 *py_states,
                                            frmc_mt19937 *np_states, frmc_run_stats *stats) {
    FRMC_REQUIRE(engines && py_states && np_states && stats && nengines >= 1, "NULL or empty argument");
    std::vector<TranslationRun> runs(nengines);
    for (int32_t r = 0; r < nengines; r++) {
        for (int32_t q = 0; q < r; q++)
            FRMC_REQUIRE(engines[q] != engines[r], "engine %d appears twice", r);
        TranslationRun &run = runs[r];
        run.e = engines[r];
        run.nsteps = nsteps;
        run.k = k;
        run.minAmp = minAmp;
        run.maxAmp = maxAmp;
        run.tolerance = tolerance;
        run.py_state = &py_states[r];
        run.np_state = &np_states[r];
        run.stats = &stats[r];
        int rc = run.init();
        if (rc) return rc;
    }
    // A few host threads share the engines (engine r to thread r % T): each
    // keeps one window in flight per engine on the engine's own stream, draws
    // and launches all its windows, then finishes them in the same order.  The
    // per-launch host cost is what bounds many small replicas, so it is spread
    // over threads; engines never share state, so the threads need no locks.
    const int32_t T = replica_threads(nengines);
    for (int32_t r = 1; r < nengines; r++)
        FRMC_REQUIRE(engines[r]->device == engines[0]->device, "replicas must share one device");
    bool batchable = g_batch_replicas != 0;
    for (int32_t r = 0; r < nengines; r++) batchable &= engines[r]->gmax == 1;
    if (batchable) {
        if (int rc = run_batched_loop(runs)) return rc;
        bool all_done = true;
        for (auto &r : runs) all_done &= r.done();
        if (!all_done) return run_batched(runs, T);
        // the accepted moves still pending, on one stream
        hipStream_t st = engines[0]->st;
        for (auto &r : runs)
            if (int rc = flush_commit(r.e, st)) return rc;
        FRMC_CHECK_HIP(hipStreamSynchronize(st));
        for (auto &r : runs)
            if (int rc = finish_windows(r.e)) return rc;
        return FRMC_OK;
    }
    std::vector<int> rcs(T, FRMC_OK);
    std::vector<std::string> errs(T);
    auto work = [&](int32_t t) {
        int rc = (int)hipSetDevice(engines[0]->device) == (int)hipSuccess ? FRMC_OK : FRMC_EHIP;
        if (rc) set_error("hipSetDevice failed in a replica thread");
        bool active = rc == FRMC_OK;
        while (active) {
            active = false;
            for (int32_t r = t; r < nengines && !rc; r += T)
                if (!runs[r].done()) {
                    rc = runs[r].launch();
                    active = true;
                }
            for (int32_t r = t; r < nengines && !rc; r += T) {
                if (runs[r].kw <= 0) continue;
                rc = runs[r].finish();
                runs[r].kw = 0;
            }
            if (rc) active = false;
        }
        for (int32_t r = t; r < nengines && !rc; r += T) rc = finish_windows(runs[r].e);
        if (rc) {
            rcs[t] = rc;
            errs[t] = frmc_last_error();
        }
    };
    std::vector<std::thread> pool;
    for (int32_t t = 1; t < T; t++) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
    for (int32_t t = 0; t < T; t++)
        if (rcs[t]) {
            set_error("%s", errs[t].c_str());
            return rcs[t];
        }
    return FRMC_OK;
}
