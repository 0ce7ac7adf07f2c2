// engine.h -- internal layout of the device-resident engine (not part of the
// C ABI).  State in HBM:
//   xyz        (n,3) float32 box coordinates, molecule / element index,
//   groups     CSR table (offsets, atom indexes),
//   per constraint c:
//     slot[c]  int32 counts [2][nel][nel][H] (+1 overflow counter): the
//              reference's hintra/hinter data, slot-exact,
//     base[c]  int32 [T][H] symmetrised pair-type sums of slot, T =
//              nel(nel+1)/2 element pairs in sorted order -- all that G(r)
//              consumes (PairDistributionConstraints.py:873-875),
//   window buffers for k proposals: int32 pair-type deltas dn[k][sum_c T*H_c]
//     (two buffers: a window accumulates into one while its evaluation
//     zeroes the other), S(Q) squared deviations, chi-square partials and
//     per-proposal completion counters.
// Window inputs (group index, old/new box coordinates of the group atoms) and
// outputs (totals, chi-squares, completion flags) live in pinned host memory
// that the kernels read and write directly, so a window needs no DMA.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "frmc_device.h"
#include "frmc_host.h"

namespace frmc {

constexpr int kMaxC = 4;   // constraints per engine
constexpr int kMaxG = 64;  // atoms per group
constexpr int kQC = 64;    // Q points per k_eval_a wave (one per lane)
constexpr int kEvalThreads = 256;  // k_eval_a block: all threads build G(r), waves 0-1 run the S(Q) sums of a leaf

struct DevC {
    int32_t kind;   // FRMC_PDC, FRMC_SFC, FRMC_PCF or FRMC_RSFC
    int32_t rspace; // 1: model on the histogram grid (PDC, PCF); 0: S(Q) (SFC, RSFC)
    Grid g;
    int32_t H, M;   // histogram bins, model points (H in r space, nQ for S(Q))
    int32_t hasW, hasSF;
    float sf;
    const float *w, *D, *shellV, *pref, *expv, *dataW;
    const float *sqm;       // S(Q) matrix transposed: [M][hpb], column q of M[r][q] zero-padded
    const float *sqi;       // the same matrix interleaved by four rows: [hpb/4][M][4], zero-padded rows
    int32_t hpb;            // H rounded up to a multiple of 32
    const int32_t *leaves;  // (start,len) pairs of NumPy's pairwise-sum leaves of length M
    const int32_t *prog;    // postfix combine program over the leaves
    int32_t nleaves, nprog;
    int32_t sameGrid;       // 1: same histogram grid as the previous constraint (bins reused)
    int32_t uOff;           // first k_eval_a work unit (r space: 1 unit, S(Q): one per pairwise-sum leaf)
    int32_t leafOff;        // S(Q): first leaf-sum word in a proposal's result block
    int32_t *base;          // [T][H]
    int32_t *slot;          // [2][nel][nel][H] + overflow counter
    int32_t dnOff;          // offset in a proposal's delta block
};

// Window staging: filled by the host in pinned memory (one block, in this
// order), copied to the device block before the window's kernels (one
// asynchronous copy on the engine stream), which read only the device copy.
struct Stage {
    int32_t *pgroup;  // [kmax]
    int32_t *poff;    // [kmax+1]
    int32_t *patoms;  // [kmax*lcap] the moved atoms of each proposal (windows of atom lists)
    float *pold;      // [kmax*lcap*3]
    float *pnew;      // [kmax*lcap*3]
    void *host = nullptr, *dev = nullptr;  // the two blocks
    size_t bytes = 0;
    int32_t *d_pgroup = nullptr, *d_poff = nullptr, *d_patoms = nullptr;
    float *d_pold = nullptr, *d_pnew = nullptr;
    bool lists = false;  // the staged window names its atoms per proposal (patoms), not by group
};

struct Prop {  // one single-atom proposal: atom, element, old and new box coordinates
    int32_t atom, el;
    float old[3], nw[3];
};

struct Pending;

struct View {
    int32_t n, nel, T, C, isPBC, ortho;
    Basis B;
    float *xyz;                    // (n,3) current box coordinates
    const int32_t *mol, *elem;
    const int32_t *goff, *gatoms;  // groups
    DevC c[kMaxC];
    int32_t dnStride;              // ints per proposal delta block
    int32_t Hmax, Mmax, maxLeaves, nunits;
    // window
    const int32_t *pgroup, *poff;  // staging (host-mapped)
    const int32_t *patoms;         // the moved atoms per proposal (at poff), or null: the group's
    const float *pold, *pnew;
    int32_t *dn;                   // delta buffer of this window
    int32_t *dn_zero;              // the other buffer, zeroed during this window's evaluation
    int64_t zero_len;
    int32_t *cnt;                  // [kmax] blocks finished per proposal (reset by the last)
    float *sqbuf;                  // [kmax][C][Mmax]
    float *chi2d;                  // [kmax][C] device partials
    unsigned long long *out_res;   // [kmax][resW] host-mapped: (seq << 32 | float bits) of the total, the
                                   // chi-squares, then the S(Q) leaf sums
    int32_t resW;                  // result words per proposal
    int32_t leafmode;              // 1: k_eval_a publishes chi-squares and leaf sums, the host combines them
    int32_t leafunits;             // S(Q) work units: 1 pairwise-sum leaves (two waves), 0 64 Q points (one wave)
    float *tse_dev;                // [kmax] totals in device memory (read by k_commit)
    unsigned long long *out_commit;  // [2] host-mapped: (seq << 32 | proposal committed by k_commit)
    uint32_t seq;
    // Spatially sorted copy of the atoms (periodic boxes): chunks of kChunkAtoms
    // atoms that lie close together, and each chunk's box in wrapped
    // fractional coordinates; k_delta skips a chunk that no pair of a
    // proposal can reach.  sx null: no sorted copy to read or maintain.
    float4 *sx;                    // [nchunk*kChunkAtoms] (x, y, z box coordinates, bits of id*64+element; -1 pad)
    float *cbox;                   // [nchunk][8] lo x,y,z, hi x,y,z (wrapped fractional), 2 unused
    const int32_t *spos;           // [n] position of each atom in sx
    int32_t nchunk;
    int32_t prune;                 // 1: this window's k_delta walks the chunks (single-atom proposals)
    float hsp[3];                  // lattice plane spacings (A) along the three box axes
    float wcut;                    // largest binned distance plus a margin for float rounding (A)
    // device window loop: the pending commit is read from here (written by
    // k_loop_step) instead of the kernel argument, and the window's
    // single-atom proposals from the table props; null otherwise
    const Pending *pend_dev;
    const Prop *props;
};

// Windows of single-atom moves travel to k_delta in the kernel arguments
// (atom index, element, old and new box coordinates per proposal, from the
// host mirror), so the kernel reads neither pinned host memory nor the
// proposal atoms' rows before its main loop.
constexpr int kInline = 64;
struct Inline {
    int32_t k;  // 0: use the staging buffers
    int32_t atom[kInline];
    int32_t el[kInline];
    float old[kInline][3];
    float nw[kInline][3];
};

// An accepted single-atom move whose commit is applied by the next window's
// k_delta (before its own deltas, which read the moved atom at its new
// position) instead of a separate k_commit launch: the old and new box
// coordinates of the atom and the proposal's delta block (base += dn).
struct Pending {
    int32_t valid;
    int32_t atom, el;
    int32_t slot;  // the atom's position in the sorted copy (-1: none)
    float old[3], nw[3];
    const int32_t *dn;
};

// Device window loop (window_loop.hip): the walk of a window's totals, the
// rewind of both random streams and the next window's draws run on the
// device (k_loop_step, one workgroup), queued with the windows' k_delta and
// k_eval launches, so the host is off the per-window critical path.
constexpr int32_t kLoopMaxK = 128;     // proposals per window
constexpr int32_t kLoopWindows = 256;  // windows per batch (the host syncs once per batch)
struct LoopMt {
    uint32_t mt[624];
    int32_t pos;
};
struct LoopAcc {  // an accepted move of the batch (host mirror)
    int32_t atom;
    float box[3], real[3];
};

// The device loop's state: both streams at the start of the window in
// flight (or of the next one), the accepted state, the batch's counters and
// the window in flight's proposals as its walk needs them.
struct LoopHead {
    float cur_tse;
    float cur_chi2[kMaxC];
    int32_t pad;
    int64_t step;        // proposals walked in the batch (log index)
    int32_t nacc, ntol;  // accepted and tolerated (accepted worse) moves of the batch
    int32_t error;       // kLoopErr*: a draw the device loop does not make
    int32_t kw;          // proposals of the window in flight (0: none)
    int32_t nwin, neval; // windows drawn and proposals drawn in the batch
    Pending pend;        // accepted move the next k_delta commits
};
static_assert(sizeof(Pending) == 48 && sizeof(LoopHead) % 4 == 0 && sizeof(LoopHead) <= 4 * 64,
              "the loop header is read as one word per lane of a wave");
struct LoopState {
    LoopMt py, np;
    LoopHead h;
    // per proposal of the window in flight (atom, element and box
    // coordinates are in the proposal table k_delta reads)
    int32_t gid[kLoopMaxK], slot[kLoopMaxK];
    int32_t py_pre[kLoopMaxK], py_post[kLoopMaxK];  // Python-stream words used before / after the acceptance draw
    float uacc[kLoopMaxK];                          // acceptance draws rounded to float32
    float nreal[kLoopMaxK][3];
};
constexpr int32_t kLoopNodes = 96;   // S(Q) leaf-combine tree nodes (all constraints) k_loop_step stages
constexpr int32_t kLoopLevels = 16;  // levels of those trees
constexpr int32_t kLoopResW = 64;    // result words per proposal

// k_loop_step's arguments: the walk of the window in flight (its totals in
// the result words), the draw of the next window, the batch's logs
struct LoopArgs {
    LoopState *L;
    int32_t reset;  // first step of a batch: the batch counters start at 0
    // walk (wk proposals; 0: no window in flight)
    int32_t wk, leafmode, resW, C, dnStride;
    float tol;
    const unsigned long long *res;
    const int32_t *win_dn;  // the walked window's delta buffer
    // NumPy's pairwise combine of the S(Q) leaf sums as a tree, nodes by
    // level: node i adds row words x and y into row word z of a proposal's
    // row (its rs words: the result words, then one per node); constraint
    // c's chi-square is row word root[c]
    const int4 *nodes;
    int32_t nnodes, nlev, rs;
    int32_t lev_off[kLoopLevels + 1];
    int32_t rspace[kMaxC], root[kMaxC];
    // draw (dk proposals; 0: none)
    int32_t dk, G, nbits;
    int32_t ident;  // group g is atom g (no group table lookups)
    float minAmp, maxAmp;
    float recip[9];
    const int32_t *goff, *gatoms, *elem, *spos;
    const float *xyz;
    float *real;
    Prop *props;  // the window's proposal table (read by the walk, rewritten by the draw)
    // logs of the batch (index = proposal walked in the batch; null: none)
    // and its accepted moves
    int32_t *glog;
    uint8_t *alog;
    float *tlog;
    LoopAcc *acc;
    // batched replicas (k_loop_step_batch): window size, the two delta
    // buffers and the one the batch's first window uses
    int32_t kr, c0;
    const int32_t *dnb[2];
    // moves left at the batch's start: a window draws at most the moves
    // left after the walks before it (the host loop's last window too)
    int64_t left;
};
int launch_loop_step(const LoopArgs &A, hipStream_t st);
// step s of a batch of M windows of R replicas (args: R device entries)
int launch_loop_step_batch(const LoopArgs *args, int32_t R, int32_t s, int32_t M, size_t shm, hipStream_t st);

constexpr int32_t kLoopErrDirection = 1;  // a zero-length direction draw (the reference would redraw)
constexpr int32_t kLoopErrWords = 2;      // a window's draws past four state blocks of the Python stream

// atoms per chunk of the sorted copy (one per thread of a k_delta block), the
// most proposals per block of the chunked k_delta, and the blocks its launch
// aims for (proposals per block = k * chunks / kDeltaBlocks, at least 1;
// 512, 1,024, 2,048 and 4,096 measured within 0.7 us of each other per
// launch on 200,000 atoms, 1,024 the fastest run)
constexpr int32_t kChunkAtoms = 256;
constexpr int32_t kPruneP = 32;
constexpr int32_t kDeltaBlocks = 1024;

// Which proposal of a window k_commit commits: p >= 0 a given one (the
// caller decided), p < 0 the first proposal that the reference's rule accepts
// (Engine.py:3308-3317) given the current total cur, the tolerance and the
// window's acceptance draws u (each rounded to float32, as the comparison
// random() > tolerance is made); none accepted: nothing is committed.
struct CommitCtl {
    int32_t k, p;
    float cur, tol;
    const int32_t *dn;  // the window's delta buffer (pair-type deltas per proposal)
    float u[kInline];
};

// proposals per k_delta block: 1 measured fastest on the 13,500-atom
// workload (k_delta 17.3 us; 2: 17.5, 4: 18.7) -- more blocks, fewer
// proposals' atomics per thread
constexpr int kDP = 1;

// Windows of many engines (replicas) launched together, one launch per kernel
// (single-atom moves).  The views are uploaded once per run; each step
// uploads the active replicas' windows, their proposals and the block maps
// of the kernels in one copy.
// other atoms per workgroup of k_delta_batch: 2,048 measured fastest on 256
// replicas (3.65M moves/s; 512: 3.10M, 1,024: 3.62M, 4,096: 3.44M,
// 8,192: 2.87M with host walks; again with walks and draws on the device:
// 5.06M against 4.66M at 1,024 and 4.85M at 4,096)
constexpr int32_t kChunk = 2048;
// proposals per workgroup of k_delta_batch (sharing each other atom's
// coordinate load): with walks and draws on the device, 256 replicas run
// 5.41M moves/s at 2 against 5.05M at 1 and 4.68M at 4 (with host walks,
// when the host bounded a step, 1 was fastest: 3.21M; 2: 3.18M; 4: 2.64M)
constexpr int32_t kBP = 2;
constexpr int32_t kBatchLdsInts = 8192;  // largest kBP delta blocks counted in LDS there (32 KB)
struct BatchWin {
    int32_t rep;             // replica (index into the views)
    int32_t k, prop;         // proposals, first of them in the proposal table
    uint32_t seq;            // the window's number (published with the results)
    int32_t *dn, *dn_zero;   // this window's delta buffer, the buffer it zeroes
    int64_t zero_len;
    Pending P;               // the replica's accepted move still to commit
};
struct BatchArgs {
    const View *views;
    const BatchWin *win;     // [active windows]
    const Prop *props;
    const int2 *map_d, *map_a, *map_b;  // per block of each kernel: (window, block within it)
};

// Kernel arguments travel by value.  HIP allows 4 KiB per launch; the largest
// launches (k_commit: View + Inline + CommitCtl; k_delta: View + Inline +
// Pending + two ints) keep a margin of more than 512 bytes for the runtime's
// hidden arguments.
static_assert(sizeof(View) + sizeof(Inline) + sizeof(CommitCtl) <= 4096 - 512, "k_commit arguments too large");
static_assert(sizeof(View) + sizeof(Inline) + sizeof(Pending) + 8 <= 4096 - 512, "k_delta arguments too large");

// launchers (engine_kernels.hip)
int launch_delta(const View &V, const Inline &I, const Pending &P, int32_t k, hipStream_t st);
int launch_eval(const View &V, int32_t k, hipStream_t st, int32_t shape);
int launch_eval_b(const View &V, int32_t k, hipStream_t st);
int launch_commit(const View &V, const Inline &I, const CommitCtl &C, hipStream_t st);
int launch_symmetrise(int32_t *base, const int32_t *slot, int32_t nel, int32_t H, hipStream_t st);
int launch_fill(int32_t *p, int64_t n, hipStream_t st);
int launch_full(const float *coords, int32_t n, const Basis &B, int isPBC, bool ortho, const int32_t *mol,
                const int32_t *elem, const HistArgs &A, int32_t *counts, hipStream_t st);
int launch_batch(const BatchArgs &A, int32_t nd, int32_t na, int32_t nb, size_t shm_d, size_t shm_a, size_t shm_b,
                 hipStream_t st);
void pairwise_plan(int32_t n, std::vector<int32_t> &leaves, std::vector<int32_t> &prog);
size_t eval_lds_bytes(const View &V);
size_t pc_lds_bytes(const View &V);  // k_eval_pc (the split-chain shapes)
size_t eval_b_lds_bytes(const View &V);

}  // namespace frmc

struct frmc_engine {
    hipStream_t st = nullptr;
    bool own_stream = false;
    int32_t n = 0, nel = 0, T = 0, C = 0, G = 0, isPBC = 0, ortho = 0, kmax = 0, gmax = 0, total_gatoms = 0;
    int32_t lcap = 0;  // atoms per proposal the staging holds: min(kMaxG, 2 gmax) (a swap moves group + partner)
    int32_t ncu = 256;
    int device = 0;                  // the HIP device the engine's memory and stream live on
    frmc::Basis B;
    float *xyz = nullptr, *dbasis = nullptr;
    int32_t *mol = nullptr, *elem = nullptr, *goff = nullptr, *gatoms = nullptr;
    std::vector<int32_t> h_goff, h_gatoms, h_elem;
    frmc::DevC c[frmc::kMaxC];
    int32_t dnStride = 0, Hmax = 0, Mmax = 0, maxLeaves = 1, nunits = 0;
    // window buffers
    int32_t *dnbuf[2] = {nullptr, nullptr};
    int cur_dn = 0;
    int32_t dn_used[2] = {0, 0};  // proposals whose deltas a buffer holds (0: all zero)
    int32_t *cnt = nullptr;
    float *sqbuf = nullptr, *chi2d = nullptr;
    frmc::Stage stage[2];
    int cur_stage = 0;     // staging buffer of the last evaluated window
    unsigned long long *h_res = nullptr;  // host-mapped results [kmax][resW]
    int32_t resW = 0;
    bool leaf_window = false;             // the last window's totals are combined on the host
    bool leaf_units = false;              // S(Q) work units are pairwise-sum leaves (else 64 Q points)
    std::vector<int32_t> h_prog[frmc::kMaxC];  // S(Q) pairwise combine programs (host copies)
    std::vector<float> h_tse, h_chi2;     // results of the last window, unpacked
    float *tse_dev = nullptr;
    unsigned long long *h_commit = nullptr;  // [2] written by k_commit (slot = window parity)
    bool device_decided = false;     // the last window's commit was queued on the device
    bool deferred = false;           // the last window's commit is left to the next k_delta
    frmc::Pending pend{};            // accepted move not yet applied on the device
    // spatially sorted copy of the atoms for k_delta's chunk walk
    float4 *sx = nullptr;
    float *cbox = nullptr;
    int32_t *spos = nullptr;
    int32_t nchunk = 0;
    bool cells = false;              // the engine keeps a sorted copy
    bool cells_stale = false;        // the copy missed commits (batched windows): rebuild before use
    int64_t commits_since_cells = 0; // accepted moves since the chunk boxes were last made tight
    std::vector<int32_t> h_spos;
    float hsp[3] = {0, 0, 0};
    float wcut = 0.f;
    const int32_t *win_dn = nullptr; // delta buffer of the last launched window
    // device window loop (allocated on first use): state, batch logs and
    // accepted moves (device, with pinned host copies), real coordinates and
    // the windows' results on the device
    frmc::LoopState *loop = nullptr, *h_loop = nullptr;
    int32_t *loop_glog = nullptr, *h_loop_glog = nullptr;
    uint8_t *loop_alog = nullptr, *h_loop_alog = nullptr;
    float *loop_tlog = nullptr, *h_loop_tlog = nullptr;
    frmc::LoopAcc *loop_acc = nullptr, *h_loop_acc = nullptr;
    frmc::Prop *loop_props = nullptr;  // the loop window's proposals
    int4 *loop_nodes = nullptr;         // S(Q) leaf-combine trees, nodes by level
    int32_t loop_nnodes = -1, loop_nlev = 0, loop_rs = 0;  // (-1: trees not built)
    int32_t loop_lev_off[frmc::kLoopLevels + 1] = {0};
    int32_t loop_root[frmc::kMaxC] = {0};
    int8_t groups_ident = -1;          // group g is atom g for every group (-1: not yet known)
    float *d_real = nullptr;
    unsigned long long *d_res = nullptr;
    int64_t loop_windows = 0;        // windows run by the device loop (all runs)
    float loop_maxabs = 0.f;         // largest box coordinate magnitude seen by the loop (bound)
    int32_t pending_check = -2;      // decision of the last device-decided window (-2: none)
    uint32_t pending_seq = 0;
    frmc::Inline last_inline;        // proposals of the last evaluated window (k = 0: staged)
    uint32_t seq = 0;
    int32_t last_k = 0;
    // host mirror of the accepted state
    float cur_tse = 0.f;
    float cur_chi2[frmc::kMaxC] = {0};
    std::vector<float> h_real, h_box;
    float recip[9] = {0};
    bool have_real = false;
    bool adaptive = true;  // window size follows the acceptance rate
    bool eval_tuned = false;     // eval_shape holds the tuned k_eval_a shapes
    int8_t pc_fit = -1;          // the split-chain shapes' LDS fits a workgroup (-1: not yet known)
    int8_t eval_shape[16] = {0};  // k_eval_a shape per window-size bucket ceil(log2 k)
    std::vector<void *> allocs, host_allocs;
    // optional per-kernel timing with HIP events on the engine stream
    bool profile = false;
    int profile_every = 1;  // time the launches of every n-th window
    struct Pending {
        int kid;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    double kernel_ms[FRMC_NKERNELS] = {0};
    int64_t kernel_launches[FRMC_NKERNELS] = {0};
};

namespace frmc {
View make_view(frmc_engine *e, int stage_idx);
int32_t eval_shape(frmc_engine *e, int32_t k);
void use_unit_layout(frmc_engine *e, bool leaf);  // S(Q) work units: pairwise leaves or 64 Q points
int launch_device_decided(frmc_engine *e, int32_t k, const int32_t *group_ids, const float *new_box,
                          const float *u, float tol);
int finish_device_decided(frmc_engine *e, int32_t k, float *tse_out);
int flush_commit(frmc_engine *e, hipStream_t on = nullptr);
int batch_prepare(frmc_engine *e, int32_t k, const int32_t *group_ids, const float *new_box, BatchWin &W,
                  Prop *props);
int settle_window(frmc_engine *e, int32_t p);
int finish_windows(frmc_engine *e);
void drop_cells(frmc_engine *e);  // the sorted copy is not maintained from here on: rebuilt before its next use
void collect_times(frmc_engine *e);
// device window loop (single-atom translation runs)
bool loop_usable(frmc_engine *e, int32_t kmax_run);
bool loop_batch_usable(frmc_engine *e, int32_t kmax_run);  // batched replicas (64-Q units, no chunk walk)
int loop_alloc(frmc_engine *e, bool single);
void loop_init_state(frmc_engine *e, LoopState &H, const frmc_mt19937 &py, const frmc_mt19937 &np);
int loop_apply(frmc_engine *e, const LoopState &H, const LoopAcc *acc, int32_t M, int64_t nlog);
int loop_begin(frmc_engine *e, const frmc_mt19937 &py, const frmc_mt19937 &np);
int loop_plan(frmc_engine *e, float maxAmp, int32_t want, int32_t &M);
int loop_batch(frmc_engine *e, int32_t k, int32_t M, float minAmp, float maxAmp, float tol);

// bracket one kernel launch with events when profiling is on
struct KTimer {
    frmc_engine *e;
    int kid;
    hipEvent_t a = nullptr;
    KTimer(frmc_engine *e_, int kid_);
    ~KTimer();
};
}  // namespace frmc
