// reciprocal.hip -- the G(r) -> S(Q) sine transform as a standalone entry
// (the transform the S(Q) constraints apply, StructureFactorConstraints.py:
// 772-773, 1255-1256) and the reciprocal_space extension
// (Extensions/reciprocal_space.pyx).
//
//   k_sq_exact   S[p][q] = sum_r G[p][r]*M[r][q] (+1), the float32 product
//                rounded, then summed sequentially over r from +0 -- the
//                order of NumPy's axis-0 sum, so bit-identical to the
//                constraint's S(Q).  One lane per (p, q); M rows stream
//                coalesced across the lanes of a wave, G broadcast from LDS.
//   k_sq_mfma    the same contraction for k curves at once on the matrix
//                cores (v_mfma_f32_16x16x4_f32, float32 inputs, float32
//                accumulation): LDS-tiled 64x64 block tiles, the r range
//                split over blocks, partials added by k_sq_splitsum.  Not
//                bit-identical to the sequential sum; its error against a
//                float64 transform is measured.
//   k_Gr_to_sq / k_gr_to_sq   reciprocal_space.pyx:89-122 / :42-80: one lane
//                per Q; each term in double (sin of the float32 product q*r,
//                then the reference's operand order) added to the float32
//                accumulator in double and rounded back, as the Cython loop's
//                Python-float arithmetic does.
//   k_sq_to_Gr   the transform of reciprocal_space.pyx:131-162 (which
//                raises for any input there), in double, rounded once.
#include <algorithm>
#include <mutex>

#include "frmc_device.h"
#include "frmc_host.h"

namespace frmc {

namespace {

constexpr int kSqT = 256;      // threads per block of the exact transform
constexpr int kGChunk = 2048;  // G(r) values staged in LDS per pass

__global__ void __launch_bounds__(kSqT) k_sq_exact(const float *__restrict__ G, int32_t k, int32_t H,
                                                   const float *__restrict__ M, int32_t nQ, int32_t addOne,
                                                   float *__restrict__ S) {
    __shared__ float gs[kGChunk];
    const int32_t p = blockIdx.y;
    const int32_t q = blockIdx.x * kSqT + threadIdx.x;
    const int32_t qc = min(q, nQ - 1);  // lanes past nQ run the chain on the last column, store nothing
    const float *g = G + (int64_t)p * H;
    float acc = 0.0f;  // NumPy's add reduction starts from +0 (a sum of -0 terms is +0)
    for (int32_t c0 = 0; c0 < H; c0 += kGChunk) {
        const int32_t len = min(kGChunk, H - c0);
        __syncthreads();
        for (int32_t i = threadIdx.x; i < len; i += kSqT) gs[i] = g[c0 + i];
        __syncthreads();
        const float *m = M + (int64_t)c0 * nQ + qc;
        int32_t i = 0;
        // batches of 16 rows: the loads issue together, then the dependent adds
        for (; i + 16 <= len; i += 16) {
            float mv[16];
#pragma unroll
            for (int j = 0; j < 16; j++) mv[j] = m[(int64_t)(i + j) * nQ];
            float pr[16];
#pragma unroll
            for (int j = 0; j < 16; j++) pr[j] = gs[i + j] * mv[j];
#pragma unroll
            for (int j = 0; j < 16; j++) acc = acc + pr[j];
        }
        for (; i < len; i++) {
            const float pr = gs[i] * m[(int64_t)i * nQ];
            acc = acc + pr;
        }
    }
    if (q < nQ) S[(int64_t)p * nQ + q] = addOne ? acc + 1.0f : acc;
}

typedef float floatx4 __attribute__((ext_vector_type(4)));

// Matrix-core transform: C (k x nQ) = G (k x H) . M (H x nQ), 64x64 output
// tile per block of four waves (each a 32x32 quarter as 2x2 MFMA tiles),
// the r range in chunks of kKC staged through LDS, the next chunk's global
// loads in flight while the current one is multiplied.  The r range is split
// over gridDim.z blocks (enough blocks to fill the chip at small k); their
// partial tiles go to a buffer that k_sq_splitsum adds in split order.
//
// v_mfma_f32_16x16x4_f32 operand layout (64 lanes): lane l supplies
// A[l%16][l/16] and B[l/16][l%16] of a 16x4 by 4x16 product and holds
// D[4*(l/16)+i][l%16], i = 0..3, of the 16x16 accumulator.
constexpr int kTP = 64, kTQ = 64, kKC = 64;
constexpr int kAP = kKC + 2;   // As row stride: fragment reads and row stores hit distinct banks
constexpr int kBP = kTQ + 16;  // Bs row stride: the two rows a half-wave reads land in disjoint banks
constexpr int kNA = kTP * kKC / 256, kNB = kKC * kTQ / 256;  // chunk elements per thread

__global__ void __launch_bounds__(256) k_sq_mfma(const float *__restrict__ G, int32_t k, int32_t H,
                                                 const float *__restrict__ M, int32_t nQ, int32_t kPerSplit,
                                                 int32_t addOne, float *__restrict__ out, int32_t direct) {
    __shared__ float As[kTP * kAP];
    __shared__ float Bs[kKC * kBP];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wp = w >> 1, wq = w & 1;
    const int32_t p0 = blockIdx.y * kTP, q0 = blockIdx.x * kTQ;
    const int32_t kb0 = blockIdx.z * kPerSplit;
    const int32_t kend = min(H, kb0 + kPerSplit);
    float ar[kNA], br[kNB];
    // chunk loads: A 64 rows (p) x kKC r, B kKC r x 64 q; indexes clamped into
    // the arrays, values past the split's r range (or past k / nQ) unused or zero
    auto load = [&](int32_t kc0) {
#pragma unroll
        for (int i = 0; i < kNA; i++) {
            const int idx = t + 256 * i;
            const int32_t ra = kc0 + idx % kKC;
            const float va = G[(int64_t)min(p0 + idx / kKC, k - 1) * H + min(ra, kend - 1)];
            ar[i] = ra < kend ? va : 0.0f;
        }
#pragma unroll
        for (int i = 0; i < kNB; i++) {
            const int idx = t + 256 * i;
            const int32_t rb = kc0 + idx / kTQ;
            const float vb = M[(int64_t)min(rb, kend - 1) * nQ + min(q0 + idx % kTQ, nQ - 1)];
            br[i] = rb < kend ? vb : 0.0f;
        }
    };
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    if (kb0 < kend) load(kb0);
    for (int32_t kc0 = kb0; kc0 < kend; kc0 += kKC) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kNA; i++) {
            const int idx = t + 256 * i;
            As[(idx / kKC) * kAP + idx % kKC] = ar[i];
        }
#pragma unroll
        for (int i = 0; i < kNB; i++) {
            const int idx = t + 256 * i;
            Bs[(idx / kTQ) * kBP + idx % kTQ] = br[i];
        }
        __syncthreads();
        if (kc0 + kKC < kend) load(kc0 + kKC);
#pragma unroll
        for (int kb = 0; kb < kKC; kb += 4) {
            float af[2], bf[2];
#pragma unroll
            for (int i = 0; i < 2; i++) af[i] = As[(wp * 32 + 16 * i + (lane & 15)) * kAP + kb + (lane >> 4)];
#pragma unroll
            for (int j = 0; j < 2; j++) bf[j] = Bs[(kb + (lane >> 4)) * kBP + wq * 32 + 16 * j + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
    }
    float *dst = direct ? out : out + (int64_t)blockIdx.z * k * nQ;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int32_t q = q0 + wq * 32 + 16 * j + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int32_t p = p0 + wp * 32 + 16 * i + 4 * (lane >> 4) + v;
                float x = acc[i][j][v];
                if (direct && addOne) x = x + 1.0f;
                if (p < k && q < nQ) dst[(int64_t)p * nQ + q] = x;
            }
        }
}

// S = sum over the splits of the partial products, in split order (+1)
__global__ void __launch_bounds__(256) k_sq_splitsum(const float *__restrict__ part, int32_t nsplit, int64_t len,
                                                     int32_t addOne, float *__restrict__ S) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
        float x = 0.0f;
        // eight independent loads in flight, then their adds in order
        for (int32_t z0 = 0; z0 < nsplit; z0 += 8) {
            float v[8];
#pragma unroll
            for (int j = 0; j < 8; j++) v[j] = part[(int64_t)min(z0 + j, nsplit - 1) * len + i];
#pragma unroll
            for (int j = 0; j < 8; j++)
                if (z0 + j < nsplit) x = x + v[j];
        }
        S[i] = addOne ? x + 1.0f : x;
    }
}

// reciprocal_space.pyx:89-122: sq = ones; sq[q] += dr*(sin(q*r)/q)*Gr[r]
// (Cython: q*r a float32 product; sin, the division and the products in
// double; the += in double, rounded to the float32 element).
__global__ void __launch_bounds__(256) k_Gr_to_sq(const float *__restrict__ dist, const float *__restrict__ Gr,
                                                  int32_t n, const float *__restrict__ qr, int32_t m,
                                                  float *__restrict__ sq) {
    __shared__ float rs[kGChunk], gs[kGChunk];
    const int32_t qi = blockIdx.x * blockDim.x + threadIdx.x;
    const float q = qr[min(qi, m - 1)];
    const float dr = load_dev(dist + 1) - load_dev(dist);
    float acc = 1.0f;
    for (int32_t c0 = 0; c0 < n; c0 += kGChunk) {
        const int32_t len = min(kGChunk, n - c0);
        __syncthreads();
        for (int32_t i = threadIdx.x; i < len; i += blockDim.x) {
            rs[i] = dist[c0 + i];
            gs[i] = Gr[c0 + i];
        }
        __syncthreads();
        for (int32_t i = 0; i < len; i++) {
            const float x = q * rs[i];
            double t = sin((double)x) / (double)q;
            t = (double)dr * t;
            t = t * (double)gs[i];
            acc = (float)((double)acc + t);
        }
    }
    if (qi < m) sq[qi] = acc;
}

// reciprocal_space.pyx:42-80: fact = 4*PI*rho (float32, PI = 3.1415927);
// sq[q] += fact*(dr*r*(sin(q*r)/q)*(gr[r]-1)) with dr*r and gr[r]-1 in
// float32 and the rest as above.
__global__ void __launch_bounds__(256) k_gr_to_sq(const float *__restrict__ dist, const float *__restrict__ gr,
                                                  int32_t n, const float *__restrict__ qr, int32_t m, float rho,
                                                  float *__restrict__ sq) {
    __shared__ float rs[kGChunk], gs[kGChunk];
    const int32_t qi = blockIdx.x * blockDim.x + threadIdx.x;
    const float q = qr[min(qi, m - 1)];
    const float dr = load_dev(dist + 1) - load_dev(dist);
    float fact = 4.0f * 3.1415927f;
    fact = fact * rho;
    float acc = 1.0f;
    for (int32_t c0 = 0; c0 < n; c0 += kGChunk) {
        const int32_t len = min(kGChunk, n - c0);
        __syncthreads();
        for (int32_t i = threadIdx.x; i < len; i += blockDim.x) {
            rs[i] = dist[c0 + i];
            gs[i] = gr[c0 + i];
        }
        __syncthreads();
        for (int32_t i = 0; i < len; i++) {
            const float r = rs[i];
            const float drr = dr * r;
            const float x = q * r;
            double t = sin((double)x) / (double)q;
            t = (double)drr * t;
            const float g1 = gs[i] - 1.0f;
            t = t * (double)g1;
            t = (double)fact * t;
            acc = (float)((double)acc + t);
        }
    }
    if (qi < m) sq[qi] = acc;
}

// G(r) = (2/PI) * sum_q Q (S(Q)-1) dq sin(Q r), dq = qValues[1]-qValues[0],
// evaluated in double and rounded once.
__global__ void __launch_bounds__(256) k_sq_to_Gr(const float *__restrict__ qv, const float *__restrict__ rv,
                                                  int32_t nr, const float *__restrict__ sq, int32_t m,
                                                  float *__restrict__ Gr) {
    __shared__ float qs[kGChunk], ss[kGChunk];
    const int32_t ri = blockIdx.x * blockDim.x + threadIdx.x;
    const double r = (double)rv[min(ri, nr - 1)];
    const double dq = (double)load_dev(qv + 1) - (double)load_dev(qv);
    double acc = 0.0;
    for (int32_t c0 = 0; c0 < m; c0 += kGChunk) {
        const int32_t len = min(kGChunk, m - c0);
        __syncthreads();
        for (int32_t i = threadIdx.x; i < len; i += blockDim.x) {
            qs[i] = qv[c0 + i];
            ss[i] = sq[c0 + i];
        }
        __syncthreads();
        for (int32_t i = 0; i < len; i++) {
            const double q = (double)qs[i];
            acc += q * ((double)ss[i] - 1.0) * sin(q * r);
        }
    }
    if (ri < nr) Gr[ri] = (float)(2.0 / 3.141592653589793 * dq * acc);
}

}  // namespace

}  // namespace frmc

using namespace frmc;

namespace {

// HIP events around the transform kernels when per-thread timing is on
// (frmc_set_timing); the stop synchronises and sets frmc_last_kernel_ms.
struct KernelTimer {
    hipEvent_t a = nullptr, b = nullptr;
    bool on = false;
    hipError_t start(hipStream_t st) {
        on = timing_enabled();
        if (!on) return hipSuccess;
        hipError_t e = hipEventCreate(&a);
        if (e == hipSuccess) e = hipEventCreate(&b);
        if (e == hipSuccess) e = hipEventRecord(a, st);
        return e;
    }
    hipError_t stop(hipStream_t st) {
        if (!on) return hipSuccess;
        float ms = 0.f;
        hipError_t e = hipEventRecord(b, st);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        set_last_kernel_ms(ms);
        return e;
    }
    ~KernelTimer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

}  // namespace

extern "C" int frmc_sq_transform(const float *Gr, int32_t k, int32_t H, const float *sqMatrix, int32_t nQ,
                                 int addOne, int mode, float *Sq, void *stream) {
    FRMC_REQUIRE(k >= 0 && H >= 1 && nQ >= 0, "sq_transform: bad sizes k=%d H=%d nQ=%d", k, H, nQ);
    FRMC_REQUIRE(mode == FRMC_SQ_EXACT || mode == FRMC_SQ_MFMA, "sq_transform: unknown mode %d", mode);
    if (k == 0 || nQ == 0) return FRMC_OK;
    FRMC_REQUIRE(Gr && sqMatrix && Sq, "sq_transform: NULL array");
    FRMC_REQUIRE((int64_t)H * nQ < ((int64_t)1 << 40), "sq_transform: matrix too large");
    hipStream_t st = as_stream(stream);
    KernelTimer timer;
    if (mode == FRMC_SQ_EXACT) {
        FRMC_REQUIRE(k <= 65535, "sq_transform: at most 65535 curves per call");
        FRMC_CHECK_HIP(timer.start(st));
        hipLaunchKernelGGL(k_sq_exact, dim3(blocks_for(nQ, kSqT), k), dim3(kSqT), 0, st, Gr, k, H, sqMatrix, nQ,
                           (int32_t)(addOne != 0), Sq);
        FRMC_CHECK_LAUNCH();
        FRMC_CHECK_HIP(timer.stop(st));
        return FRMC_OK;
    }
    const int64_t bq = (nQ + kTQ - 1) / kTQ, bp = (k + kTP - 1) / kTP;
    FRMC_REQUIRE(bp <= 65535, "sq_transform: too many curves per call");
    // split r so that about 1024 blocks run (at least one chunk each)
    const int64_t chunks = (H + kKC - 1) / kKC;
    int64_t nsplit = std::max<int64_t>(1, std::min<int64_t>((1024 + bq * bp - 1) / (bq * bp), chunks));
    const int32_t per = (int32_t)(((chunks + nsplit - 1) / nsplit) * kKC);
    nsplit = (H + per - 1) / per;
    float *part = nullptr;
    if (nsplit > 1) {
        hipError_t e = pool_malloc((void **)&part, sizeof(float) * nsplit * k * nQ, st);
        if (e != hipSuccess) {
            set_error("hipMallocAsync of the split partials failed: %s", hipGetErrorString(e));
            return FRMC_ENOMEM;
        }
    }
    hipError_t le = timer.start(st);
    if (le == hipSuccess) {
        hipLaunchKernelGGL(k_sq_mfma, dim3(bq, bp, nsplit), dim3(256), 0, st, Gr, k, H, sqMatrix, nQ, per,
                           (int32_t)(addOne != 0), nsplit > 1 ? part : Sq, (int32_t)(nsplit == 1));
        le = hipGetLastError();
    }
    if (le == hipSuccess && nsplit > 1) {
        const int64_t len = (int64_t)k * nQ;
        hipLaunchKernelGGL(k_sq_splitsum, dim3((unsigned)std::min<int64_t>(blocks_for(len, 256), 4096)), dim3(256),
                           0, st, part, (int32_t)nsplit, len, (int32_t)(addOne != 0), Sq);
        le = hipGetLastError();
    }
    if (le == hipSuccess) le = timer.stop(st);
    if (part) (void)hipFreeAsync(part, st);
    if (le != hipSuccess) {
        set_error("sq_transform failed: %s", hipGetErrorString(le));
        return FRMC_EHIP;
    }
    return FRMC_OK;
}

// ------------------------------------------------- no-allocation variant
//
// The matrix-core transform's split partials in a file-scope device array
// instead of stream-ordered pool memory, for callers that cannot allocate
// (code under HIP graph capture, allocation-free serving loops).  Each GPU
// the library runs on holds its own copy (16 MiB) until the process ends.
// Every call on a GPU shares that GPU's copy, and the next call overwrites
// it: calls that overlap on one GPU (other streams, other threads) race.
// k_sq_mfma stores every partial element it owns, so nothing needs clearing.
constexpr int64_t kSqScratchFloats = (int64_t)1 << 22;
static __device__ __align__(256) float scratch[kSqScratchFloats];

static std::mutex g_sq_scratch_mu;
static float *g_sq_scratch[64] = {nullptr};  // device address of scratch, per device

// This device's copy of scratch: its address is taken once per device
// (with that device current), then reused; host code never uses the
// symbol itself, which is only a host-side stand-in.
static int sq_scratch_of_current_device(float **out) {
    int dev = 0;
    FRMC_CHECK_HIP(hipGetDevice(&dev));
    FRMC_REQUIRE(dev >= 0 && dev < 64, "sq_transform_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_sq_scratch_mu);
    if (!g_sq_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("sq_transform_noalloc: hipGetSymbolAddress of the scratch array failed: %s",
                      hipGetErrorString(e));
            return FRMC_EHIP;
        }
        g_sq_scratch[dev] = (float *)p;
    }
    *out = g_sq_scratch[dev];
    return FRMC_OK;
}

extern "C" int frmc_sq_transform_noalloc(const float *Gr, int32_t k, int32_t H, const float *sqMatrix, int32_t nQ,
                                         int addOne, int mode, float *Sq, void *stream) {
    FRMC_REQUIRE(k >= 0 && H >= 1 && nQ >= 0, "sq_transform_noalloc: bad sizes k=%d H=%d nQ=%d", k, H, nQ);
    FRMC_REQUIRE(mode == FRMC_SQ_EXACT || mode == FRMC_SQ_MFMA, "sq_transform_noalloc: unknown mode %d", mode);
    if (k == 0 || nQ == 0) return FRMC_OK;
    FRMC_REQUIRE(Gr && sqMatrix && Sq, "sq_transform_noalloc: NULL array");
    FRMC_REQUIRE((int64_t)H * nQ < ((int64_t)1 << 40), "sq_transform_noalloc: matrix too large");
    hipStream_t st = as_stream(stream);
    if (mode == FRMC_SQ_EXACT) {
        // no scratch: the exact kernel of frmc_sq_transform
        FRMC_REQUIRE(k <= 65535, "sq_transform_noalloc: at most 65535 curves per call");
        hipLaunchKernelGGL(k_sq_exact, dim3(blocks_for(nQ, kSqT), k), dim3(kSqT), 0, st, Gr, k, H, sqMatrix, nQ,
                           (int32_t)(addOne != 0), Sq);
        FRMC_CHECK_LAUNCH();
        return FRMC_OK;
    }
    // the same r split as frmc_sq_transform, so the results are identical
    const int64_t bq = (nQ + kTQ - 1) / kTQ, bp = (k + kTP - 1) / kTP;
    FRMC_REQUIRE(bp <= 65535, "sq_transform_noalloc: too many curves per call");
    const int64_t chunks = (H + kKC - 1) / kKC;
    int64_t nsplit = std::max<int64_t>(1, std::min<int64_t>((1024 + bq * bp - 1) / (bq * bp), chunks));
    const int32_t per = (int32_t)(((chunks + nsplit - 1) / nsplit) * kKC);
    nsplit = (H + per - 1) / per;
    float *part = nullptr;
    if (nsplit > 1) {
        const int64_t need = nsplit * (int64_t)k * nQ;
        FRMC_REQUIRE(need <= kSqScratchFloats,
                     "sq_transform_noalloc: k=%d, H=%d, nQ=%d need %lld floats of split partials; the "
                     "no-allocation entry holds at most %lld (16 MiB): use frmc_sq_transform",
                     k, H, nQ, (long long)need, (long long)kSqScratchFloats);
        if (int rc = sq_scratch_of_current_device(&part)) return rc;
    }
    hipLaunchKernelGGL(k_sq_mfma, dim3(bq, bp, nsplit), dim3(256), 0, st, Gr, k, H, sqMatrix, nQ, per,
                       (int32_t)(addOne != 0), nsplit > 1 ? part : Sq, (int32_t)(nsplit == 1));
    FRMC_CHECK_LAUNCH();
    if (nsplit > 1) {
        const int64_t len = (int64_t)k * nQ;
        hipLaunchKernelGGL(k_sq_splitsum, dim3((unsigned)std::min<int64_t>(blocks_for(len, 256), 4096)), dim3(256),
                           0, st, part, (int32_t)nsplit, len, (int32_t)(addOne != 0), Sq);
        FRMC_CHECK_LAUNCH();
    }
    return FRMC_OK;
}

extern "C" int frmc_Gr_to_sq(const float *distances, const float *Gr, int32_t n, const float *qrange, int32_t m,
                             float *sq, void *stream) {
    FRMC_REQUIRE(n >= 2 && m >= 0, "Gr_to_sq: need at least two distances (n=%d)", n);
    if (m == 0) return FRMC_OK;
    FRMC_REQUIRE(distances && Gr && qrange && sq, "Gr_to_sq: NULL array");
    hipLaunchKernelGGL(k_Gr_to_sq, dim3(blocks_for(m, 256)), dim3(256), 0, as_stream(stream), distances, Gr, n,
                       qrange, m, sq);
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

extern "C" int frmc_gr_to_sq(const float *distances, const float *gr, int32_t n, const float *qrange, int32_t m,
                             float rho, float *sq, void *stream) {
    FRMC_REQUIRE(n >= 2 && m >= 0, "gr_to_sq: need at least two distances (n=%d)", n);
    if (m == 0) return FRMC_OK;
    FRMC_REQUIRE(distances && gr && qrange && sq, "gr_to_sq: NULL array");
    hipLaunchKernelGGL(k_gr_to_sq, dim3(blocks_for(m, 256)), dim3(256), 0, as_stream(stream), distances, gr, n,
                       qrange, m, rho, sq);
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}

extern "C" int frmc_sq_to_Gr(const float *qValues, const float *rValues, int32_t nr, const float *sq, int32_t m,
                             float *Gr, void *stream) {
    FRMC_REQUIRE(m >= 2 && nr >= 0, "sq_to_Gr: need at least two Q values (m=%d)", m);
    if (nr == 0) return FRMC_OK;
    FRMC_REQUIRE(qValues && rValues && sq && Gr, "sq_to_Gr: NULL array");
    hipLaunchKernelGGL(k_sq_to_Gr, dim3(blocks_for(nr, 256)), dim3(256), 0, as_stream(stream), qValues, rValues,
                       nr, sq, m, Gr);
    FRMC_CHECK_LAUNCH();
    return FRMC_OK;
}
