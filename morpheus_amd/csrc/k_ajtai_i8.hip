// Ajtai-style linear commitment with an int8 public matrix on INT8 MFMA:
//   c_i = sum_j M_ij z_j  (mod p),   i < rows, j < n,
// M_ij in [-128, 127] expanded from a 32-byte key with ChaCha8, z_j Goldilocks words.
//
// This is the benchmark's third configuration ("witness of 2^22 elements decomposed into
// 8 int8 limbs, 256-row int8 matrix from a ChaCha8 seed, INT8 MFMA with int32 accumulation,
// exact recombination and reduction mod p"). The reference has no function with these
// semantics: its Ajtai is a ring-SIS commitment over almost-Goldilocks
// (cuda_almost_goldilocks/ajtai.cuh:94-275); its INT8 tensor-core paths are a multifold
// GEMM (ajtai_tc.cuh:166-225) and a throughput probe (ajtai_tc_commit_probe.cuh:38-99).
// The ChaCha8 block is the reference's (ajtai_chacha8.cuh:60-106). Definitions fixed here:
//   M_ij = (int8) byte (j mod 64) of ChaCha8(key, counter = j / 64, nonce = (i, 0, 0))
//   limbs: z_j = sum_l 2^(8l) d_jl with bytes d_jl; signed operands d'_jl = d_jl - 128
//   c_i   = sum_l 2^(8l) (sum_j M_ij d'_jl + 128 sum_j M_ij)   (mod p)   = M z (mod p)
//
// MI355X layout and kernels:
// * M is expanded once (setup) into HBM pre-tiled in MFMA A-fragment order for
//   v_mfma_i32_16x16x64_i8: for every 64-column k-step and 16-row tile, the 64 lanes'
//   16-byte fragments are 1 KiB contiguous (lane l = r + 16 g holds row r, columns
//   16 g .. 16 g + 15 of the step), so every operand load is a coalesced 16-byte load.
//   Row sums are accumulated during expansion.
// * The witnesses are split into 8 limb planes while a workgroup stages them into LDS, in
//   the same fragment order (lane l = c + 16 g holds plane c of a 16-plane tile, columns
//   16 g .. 16 g + 15). A and B use the same (lane group, byte) -> column assignment, so the
//   MFMA's own k order is irrelevant.
// * The GEMM: one workgroup per K-chunk covers all rows and all planes; it stages DEPTH
//   k-steps of A and of the limb planes in LDS (double-buffered: the next group's global
//   loads overlap this group's MFMAs), so every HBM byte is read once; each wave owns an
//   RTW x CTW block of 16x16 accumulator tiles. |chunk sum| <= 2^14 * chunk <= 2^28, exact in
//   int32; chunk partials are summed in int64 and recombined mod p by the finish kernel.
//   At B = 1 the kernel streams the 1 GiB matrix once (HBM-bound).
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "chacha.h"
#include "common.h"
#include "profile.h"
#include "sync.h"

typedef int v4i __attribute__((ext_vector_type(4)));

static constexpr int AJ_MAX_BATCH = 32;

// byte offset of entry (i, j) of a tiled operand with `tiles` 16-row tiles per k-step
__host__ __device__ __forceinline__ size_t aj_offset(size_t i, size_t j, size_t tiles) {
    const size_t kt = j >> 6, g = (j >> 4) & 3, e = j & 15, t = i >> 4, r = i & 15;
    return ((kt * tiles + t) * 64 + r + 16 * g) * 16 + e;
}

// One thread per (ChaCha block of 64 columns, row); rows fastest so stores coalesce.
__global__ void k_aj_expand(const u32 *__restrict__ key_words, int rows, size_t n, uint4 *__restrict__ mt,
                            unsigned long long *__restrict__ rowsum) {
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= (size_t)rows * (n / 64)) return;
    const u32 i = (u32)(t % rows);
    const u32 b = (u32)(t / rows);
    u32 key[8];
#pragma unroll
    for (int k = 0; k < 8; k++) key[k] = key_words[k];
    const u32 nonce[3] = {i, 0, 0};
    u32 w[16];
    chacha::block<8>(key, b, nonce, w);
    const int RT = rows / 16;
    long long sum = 0;
#pragma unroll
    for (int g = 0; g < 4; g++) {
        mt[aj_offset(i, (size_t)b * 64 + 16 * g, RT) / 16] = make_uint4(w[4 * g], w[4 * g + 1], w[4 * g + 2], w[4 * g + 3]);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const u32 v = w[4 * g + k];
            sum += (signed char)v + (signed char)(v >> 8) + (signed char)(v >> 16) + (signed char)(v >> 24);
        }
    }
    atomicAdd(&rowsum[i], (unsigned long long)sum);
}

// limb[l][q] = bytes l of w[0..3] packed little-endian (a 4 x 8 byte transpose, 16 v_perm).
// v_perm_b32(hi, lo, sel): result byte k = byte sel_k of the 8-byte value hi:lo.
__device__ __forceinline__ void aj_transpose4(const u64 w[4], u32 limb[8][4], int q) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const u32 x0 = (u32)(w[0] >> (32 * h)), x1 = (u32)(w[1] >> (32 * h));
        const u32 x2 = (u32)(w[2] >> (32 * h)), x3 = (u32)(w[3] >> (32 * h));
        const u32 a = __builtin_amdgcn_perm(x1, x0, 0x05010400u);  // x0.b0 x1.b0 x0.b1 x1.b1
        const u32 b = __builtin_amdgcn_perm(x1, x0, 0x07030602u);  // x0.b2 x1.b2 x0.b3 x1.b3
        const u32 c = __builtin_amdgcn_perm(x3, x2, 0x05010400u);
        const u32 d = __builtin_amdgcn_perm(x3, x2, 0x07030602u);
        limb[4 * h + 0][q] = __builtin_amdgcn_perm(c, a, 0x05040100u);
        limb[4 * h + 1][q] = __builtin_amdgcn_perm(c, a, 0x07060302u);
        limb[4 * h + 2][q] = __builtin_amdgcn_perm(d, b, 0x05040100u);
        limb[4 * h + 3][q] = __builtin_amdgcn_perm(d, b, 0x07060302u);
    }
}

// grid: one workgroup per K-chunk of `kchunk` columns; block: 512 threads; wave
// (wr, wc) < (WR, CT / CTW) owns row tiles wr*RTW .. +RTW-1 (< RT) and plane tiles
// wc*CTW .. +CTW-1. partial[(chunk * rows + row) * (CTT * 16) + col] = chunk sum (int32).
static constexpr int AJF_THREADS = 512;

template <int RTW, int CTW, int CT, int DEPTH>
__global__ __launch_bounds__(AJF_THREADS) void k_aj_gemm(const v4i *__restrict__ mt, const u64 *__restrict__ z,
                                                               int RT, int batch, int ones, int WR, size_t n,
                                                               size_t kchunk, int *__restrict__ partial) {
    constexpr int TILES = 16 + CT;
    constexpr int A_ITEMS = DEPTH * 16 * 64 / AJF_THREADS;  // v4i per thread (full RT)
    __shared__ v4i lds[2][DEPTH][TILES][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = wave % WR, wc = wave / WR;
    const bool computes = wave < WR * (CT / CTW);
    const int lc = lane & 15, q = lane >> 4;
    const int CTT = (8 * batch + ones + 15) / 16;
    const size_t steps = n >> 6;
    const size_t kt0 = (size_t)blockIdx.x * (kchunk >> 6);
    size_t kt1 = kt0 + (kchunk >> 6);
    if (kt1 > steps) kt1 = steps;

    // staging registers: A fragments, and one (witness, 16-column group) of words per thread
    v4i ra[A_ITEMS];
    u64 rz[16];
    const int a_per_step = RT * 64;
    const int wb = batch + ones;  // witnesses, plus the all-ones plane when `ones`
    const int b_items = DEPTH * wb * 4;
    auto stage_load = [&](size_t kt) {
#pragma unroll
        for (int m = 0; m < A_ITEMS; m++) {
            const int e = threadIdx.x + m * AJF_THREADS;
            const int d = e / (16 * 64), r = e % (16 * 64);
            ra[m] = v4i{0, 0, 0, 0};
            // the matrix is read exactly once per commit: non-temporal loads keep it from
            // displacing anything in the caches (7.2 vs 5.3-5.7 TB/s on a plain 1 GiB read)
            if (r < a_per_step && kt + d < kt1) ra[m] = __builtin_nontemporal_load(&mt[(kt + d) * a_per_step + r]);
        }
        const int e = threadIdx.x;
        if (e < b_items) {
            const int d = e / (wb * 4), r = e % (wb * 4);
            if (kt + d < kt1 && (r >> 2) < batch) {
                const u64 *src = z + (size_t)(r >> 2) * n + (kt + d) * 64 + (r & 3) * 16;
#pragma unroll
                for (int k = 0; k < 16; k++) rz[k] = src[k];
            }
        }
    };
    auto stage_store = [&](int buf, size_t kt) {
#pragma unroll
        for (int m = 0; m < A_ITEMS; m++) {
            const int e = threadIdx.x + m * AJF_THREADS;
            const int d = e / (16 * 64), r = e % (16 * 64);
            if (r < a_per_step) lds[buf][d][r >> 6][r & 63] = ra[m];
        }
        const int e = threadIdx.x;
        if (e < b_items) {
            const int d = e / (wb * 4), r = e % (wb * 4);
            if (kt + d < kt1 && (r >> 2) == batch) {
                // the all-ones plane: its GEMM column is the row sum over the committed columns
                const int p = 8 * batch, g = r & 3;
                lds[buf][d][16 + (p >> 4)][(p & 15) + 16 * g] = v4i{0x01010101, 0x01010101, 0x01010101, 0x01010101};
            } else if (kt + d < kt1) {
                const int w = r >> 2, g = r & 3;
#pragma unroll
                for (int k = 0; k < 16; k++) rz[k] = gl_canon(rz[k]);
                u32 limb[8][4];  // limb[l][q] = byte l of words 4q .. 4q + 3
#pragma unroll
                for (int qq = 0; qq < 4; qq++) aj_transpose4(&rz[4 * qq], limb, qq);
#pragma unroll
                for (int l = 0; l < 8; l++) {
                    v4i o;
#pragma unroll
                    for (int qq = 0; qq < 4; qq++) o[qq] = (int)(limb[l][qq] ^ 0x80808080u);
                    const int p = 8 * w + l;
                    lds[buf][d][16 + (p >> 4)][(p & 15) + 16 * g] = o;
                }
            }
        }
    };

    v4i acc[RTW][CTW];
#pragma unroll
    for (int i = 0; i < RTW; i++)
#pragma unroll
        for (int j = 0; j < CTW; j++) acc[i][j] = v4i{0, 0, 0, 0};

    stage_load(kt0);
    stage_store(0, kt0);
    mam_barrier();
    int cur = 0;
    for (size_t kt = kt0; kt < kt1; kt += DEPTH) {
        const bool more = kt + DEPTH < kt1;
        if (more) stage_load(kt + DEPTH);
        if (computes) {
#pragma unroll
            for (int d = 0; d < DEPTH; d++) {
                v4i a[RTW];
#pragma unroll
                for (int i = 0; i < RTW; i++) a[i] = lds[cur][d][min(wr * RTW + i, 15)][lane];
#pragma unroll
                for (int j = 0; j < CTW; j++) {
                    const v4i b = lds[cur][d][16 + wc * CTW + j][lane];
#pragma unroll
                    for (int i = 0; i < RTW; i++)
                        acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[i], b, acc[i][j], 0, 0, 0);
                }
            }
        }
        if (more) stage_store(cur ^ 1, kt + DEPTH);
        mam_barrier();
        cur ^= 1;
    }
    if (!computes) return;
    const int rows = RT * 16, cols = CTT * 16;
#pragma unroll
    for (int i = 0; i < RTW; i++) {
        if (wr * RTW + i >= RT) continue;
#pragma unroll
        for (int j = 0; j < CTW; j++) {
            const int col = (wc * CTW + j) * 16 + lc;
            if (col >= cols) continue;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = (wr * RTW + i) * 16 + q * 4 + v;
                partial[((size_t)blockIdx.x * rows + row) * cols + col] = acc[i][j][v];
            }
        }
    }
}

// out[w * rows + i] = sum_l 2^(8l) (S'_{i,(w,l)} + 128 R_i) mod p. One workgroup of 256
// threads per (i, w): thread (slice s, limb l) sums chunks s, s + 32, ...; LDS reduce.
// With ones_col >= 0 the row sums come from that partial column (the all-ones plane over the
// committed columns) instead of the set-up's full-width row sums.
__global__ __launch_bounds__(256) void k_aj_finish(const int *__restrict__ partial,
                                                   const unsigned long long *__restrict__ rowsum, int rows, int cols,
                                                   int chunks, int ones_col, u64 *__restrict__ out) {
    __shared__ long long red[256];
    const int i = blockIdx.x % rows, w = blockIdx.x / rows;
    const int l = threadIdx.x & 7, sl = threadIdx.x >> 3;
    long long s = 0;
    for (int c = sl; c < chunks; c += 32) s += partial[((size_t)c * rows + i) * cols + w * 8 + l];
    red[threadIdx.x] = s;
    mam_barrier();
    for (int h = 128; h >= 8; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        mam_barrier();
    }
    if (threadIdx.x == 0) {
        long long R = 0;
        if (ones_col >= 0) {
            for (int c = 0; c < chunks; c++) R += partial[((size_t)c * rows + i) * cols + ones_col];
        } else {
            R = (long long)rowsum[i];
        }
        u64 acc = 0;
        for (int k = 0; k < 8; k++) {
            const long long v = red[k] + 128 * R;  // |v| < 2^40 < p
            const u64 r = v >= 0 ? (u64)v : GL_P - (u64)(-v);
            acc = gl_add(acc, gl_mul(r, (u64)1 << (8 * k)));
        }
        out[(size_t)w * rows + i] = acc;
    }
}

namespace {
struct AjtaiI8 {
    int rows = 0;
    size_t n = 0;
    u64 *mt = nullptr;                     // rows * n bytes, tiled
    unsigned long long *rowsum = nullptr;  // rows
};

// Chunk partials (chunks * rows * CTT * 16 ints), grown on demand. They belong to the calling
// thread, whose commits all run on that thread's stream, so one set-up serves commits from
// several threads at once (the proving service's workers share one set-up).
struct AjScratch {
    int *partial = nullptr;
    size_t bytes = 0;
};
thread_local AjScratch t_aj_scratch;

// Grow a device buffer; the stream is drained before an old buffer is freed.
int aj_reserve(void **buf, size_t *have, size_t need, hipStream_t s) {
    if (need <= *have) return 0;
    if (*buf) {
        (void)hipStreamSynchronize(s);
        (void)hipFree(*buf);
        *buf = nullptr;
        *have = 0;
    }
    if (hipMalloc(buf, need) != hipSuccess) return -1;
    *have = need;
    return 0;
}

template <int RTW, int CTW, int CT, int DEPTH>
void aj_launch(const AjtaiI8 *A, const u64 *d_z, int batch, int ones, size_t n_used, size_t kchunk, unsigned grid,
               int *partial, hipStream_t s) {
    static_assert(DEPTH * 33 * 4 <= AJF_THREADS, "one witness group per staging thread");
    static_assert((16 / RTW) * (CT / CTW) <= AJF_THREADS / 64, "wave blocks fit the workgroup");
    const int RT = A->rows / 16;
    const int WR = (RT + RTW - 1) / RTW;
    k_aj_gemm<RTW, CTW, CT, DEPTH><<<grid, AJF_THREADS, 0, s>>>((const v4i *)A->mt, d_z, RT, batch, ones, WR,
                                                                     n_used, kchunk, partial);
}
}  // namespace

// Chunk partials of the allocation-free commit (mam_ajtai_i8_commit_noalloc): one array per
// GPU, held until the process ends. Host code never passes the array's name to a kernel (on
// the host it names a stand-in, not device memory): the device address comes from
// hipGetSymbolAddress, once per device, with that device current, and is kept in
// g_scratch_dev[device].
static constexpr size_t AJ_SCRATCH_BYTES = (size_t)16 << 20;
__device__ __align__(256) int scratch[AJ_SCRATCH_BYTES / sizeof(int)];

namespace {
constexpr int AJ_SCRATCH_DEVS = 64;
std::mutex g_scratch_mu;
int *g_scratch_dev[AJ_SCRATCH_DEVS] = {};

// The current device's address of `scratch`, resolved on the first call on that device;
// nullptr when the runtime cannot resolve it.
int *scratch_on_current_device() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= AJ_SCRATCH_DEVS) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        g_scratch_dev[dev] = (int *)ptr;
    }
    return g_scratch_dev[dev];
}
}  // namespace

// Expand the matrix for (key, rows, n): rows a multiple of 16 (<= 256), n a multiple of 64
// (<= 2^30, which keeps every recombined limb sum below p).
MAM_API int mam_ajtai_i8_setup(const uint8_t *key32, int rows, uint64_t n, void **handle) {
    if (!handle || !key32 || rows <= 0 || rows % 16 || rows > 256 || n == 0 || n % 64 || n > ((uint64_t)1 << 30))
        return -2;
    *handle = nullptr;
    // set-up also resolves this device's address of the allocation-free commit's scratch
    if (!scratch_on_current_device()) return -1;
    AjtaiI8 *A = new AjtaiI8();
    A->rows = rows;
    A->n = n;
    hipStream_t s = mam_stream();
    u32 *d_key = nullptr;
    int rc = 0;
    if (hipMalloc((void **)&A->mt, (size_t)rows * n) != hipSuccess ||
        hipMalloc((void **)&A->rowsum, rows * sizeof(u64)) != hipSuccess ||
        hipMalloc((void **)&d_key, 32) != hipSuccess)
        rc = -1;
    u32 kw[8];
    for (int k = 0; k < 8; k++)
        kw[k] = (u32)key32[4 * k] | ((u32)key32[4 * k + 1] << 8) | ((u32)key32[4 * k + 2] << 16) |
                ((u32)key32[4 * k + 3] << 24);
    if (!rc && hipMemcpyAsync(d_key, kw, 32, hipMemcpyHostToDevice, s) != hipSuccess) rc = -1;
    if (!rc && hipMemsetAsync(A->rowsum, 0, rows * sizeof(u64), s) != hipSuccess) rc = -1;
    if (!rc) {
        const size_t threads = (size_t)rows * (n / 64);
        k_aj_expand<<<grid_1d(threads), MAM_BLOCK, 0, s>>>(d_key, rows, n, (uint4 *)A->mt, A->rowsum);
        if (hipGetLastError() != hipSuccess) rc = -1;
    }
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = -1;
    if (d_key) (void)hipFree(d_key);
    if (rc) {
        if (A->mt) (void)hipFree(A->mt);
        if (A->rowsum) (void)hipFree(A->rowsum);
        delete A;
        return rc;
    }
    *handle = A;
    return 0;
}

// Launch shape of one commit: plane tiles, columns per workgroup, chunk count and the bytes
// of chunk partials the GEMM writes (and k_aj_finish reads back).
struct AjPlan {
    int ones = 0, CTT = 0, CT = 0;
    size_t kchunk = 0, chunks = 0, partial_bytes = 0;
};

// Commit `batch` witnesses of n_used words each (back to back, device) against the first
// n_used matrix columns; d_out: batch * rows words (device). n_used = n uses the set-up's row
// sums; a shorter prefix (a multiple of 64, batch <= 16) adds an all-ones plane to the GEMM,
// whose column is the prefix row sum (the matrix entries depend only on (key, row, column),
// so a prefix commitment equals a commitment with a set-up of n_used columns).
static int aj_plan(const AjtaiI8 *A, const u64 *d_z, int batch, size_t n_used, const u64 *d_out, AjPlan &P) {
    if (!A || !d_z || !d_out || batch < 1 || batch > AJ_MAX_BATCH) return -2;
    if (n_used == 0 || n_used % 64 || n_used > A->n) return -2;
    P.ones = n_used != A->n ? 1 : 0;
    if (P.ones && batch > 16) return -2;
    const int nplanes = 8 * batch + P.ones;
    P.CTT = (nplanes + 15) / 16;
    P.CT = P.CTT <= 1 ? 1 : P.CTT <= 2 ? 2 : P.CTT <= 4 ? 4 : P.CTT <= 8 ? 8 : 16;
    // 2^13 columns per workgroup (512 workgroups at n = 2^22) for up to 32 planes, 2^14 for
    // wider batches (fewer partials to write and add: B = 8 0.349 -> 0.337 ms at 2^22; 2^12
    // was slower at every batch); both keep chunk sums below 2^28.
    P.kchunk = P.CT <= 2 ? ((size_t)1 << 13) : ((size_t)1 << 14);
    // MAM_AJ_KCHUNK_LOG2 (12..14) overrides the columns per workgroup
    static const int kc_log2 = [] {
        const char *e = getenv("MAM_AJ_KCHUNK_LOG2");
        const int v = e ? atoi(e) : 0;
        return (v >= 12 && v <= 14) ? v : 0;
    }();
    if (kc_log2) P.kchunk = (size_t)1 << kc_log2;
    P.chunks = (n_used + P.kchunk - 1) / P.kchunk;
    P.partial_bytes = P.chunks * A->rows * P.CTT * 16 * sizeof(int);
    return 0;
}

// Queues the GEMM and the finish on stream s, with the chunk partials at `partial`
// (P.partial_bytes, every word written by the GEMM before the finish reads it).
static int aj_enqueue(const AjtaiI8 *A, const u64 *d_z, int batch, size_t n_used, const AjPlan &P, int *partial,
                      u64 *d_out, hipStream_t s) {
    const unsigned grid = (unsigned)P.chunks;
    switch (P.CT) {
        case 1: aj_launch<2, 1, 1, 2>(A, d_z, batch, P.ones, n_used, P.kchunk, grid, partial, s); break;
        case 2: aj_launch<2, 2, 2, 2>(A, d_z, batch, P.ones, n_used, P.kchunk, grid, partial, s); break;
        case 4: aj_launch<4, 2, 4, 2>(A, d_z, batch, P.ones, n_used, P.kchunk, grid, partial, s); break;
        case 8: aj_launch<4, 4, 8, 2>(A, d_z, batch, P.ones, n_used, P.kchunk, grid, partial, s); break;
        default: aj_launch<4, 8, 16, 2>(A, d_z, batch, P.ones, n_used, P.kchunk, grid, partial, s); break;
    }
    if (hipGetLastError() != hipSuccess) return -1;
    k_aj_finish<<<batch * A->rows, 256, 0, s>>>(partial, A->rowsum, A->rows, P.CTT * 16, (int)P.chunks,
                                               P.ones ? 8 * batch : -1, d_out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

static int aj_commit(AjtaiI8 *A, const u64 *d_z, int batch, size_t n_used, u64 *d_out) {
    AjPlan P;
    int rc = aj_plan(A, d_z, batch, n_used, d_out, P);
    if (rc) return rc;
    hipStream_t s = mam_stream();
    AjScratch &sc = t_aj_scratch;
    if (aj_reserve((void **)&sc.partial, &sc.bytes, P.partial_bytes, s)) return -1;
    {
        ProfScope prof(PROF_AJTAI_GEMM, (double)A->rows * n_used + 8.0 * batch * n_used, s);
        rc = aj_enqueue(A, d_z, batch, n_used, P, sc.partial, d_out, s);
    }
    if (rc) return rc;
    return mam_sync_status();
}

// d_z: batch witnesses of n words back to back (device); d_out: batch * rows words (device).
MAM_API int mam_ajtai_i8_commit(void *handle, const u64 *d_z, int batch, u64 *d_out) {
    AjtaiI8 *A = (AjtaiI8 *)handle;
    return A ? aj_commit(A, d_z, batch, A->n, d_out) : -2;
}

MAM_API int mam_ajtai_i8_commit_prefix(void *handle, const u64 *d_z, int batch, uint64_t n_used, u64 *d_out) {
    return aj_commit((AjtaiI8 *)handle, d_z, batch, (size_t)n_used, d_out);
}

// The commit of mam_ajtai_i8_commit_prefix without any allocation: the chunk partials go to
// the file-scope `scratch` array of the current device (AJ_SCRATCH_BYTES) instead of a
// buffer the library allocates. Returns once the work is queued on the calling thread's
// stream (d_out is complete when that stream is synchronized); -2 for the arguments
// mam_ajtai_i8_commit_prefix rejects and when the partials need more than AJ_SCRATCH_BYTES
// (chunks x rows x 16 ceil((8 batch + [n_used < n]) / 16) x 4 bytes, with 2^13 columns per
// chunk up to 32 planes and 2^14 above: at rows = 256 and n_used = 2^22, batch <= 8). Every
// call on one GPU shares that GPU's array: calls that overlap on one GPU (other threads'
// streams, other caller streams) race; calls queued on one stream do not.
MAM_API int mam_ajtai_i8_commit_noalloc(void *handle, const u64 *d_z, int batch, uint64_t n_used, u64 *d_out) {
    AjtaiI8 *A = (AjtaiI8 *)handle;
    AjPlan P;
    int rc = aj_plan(A, d_z, batch, (size_t)n_used, d_out, P);
    if (rc) return rc;
    if (P.partial_bytes > AJ_SCRATCH_BYTES) return -2;
    int *ptr = scratch_on_current_device();
    if (!ptr) return -1;
    return aj_enqueue(A, d_z, batch, (size_t)n_used, P, ptr, d_out, mam_stream());
}

// Copy `count` matrix entries of one row starting at column col0 to host (int8).
// Gathers 8-entry groups g0 .. g0 + groups - 1 of one row from the tiled matrix.
__global__ void k_aj_row(const uint8_t *__restrict__ mt, int row, int RT, uint64_t g0, uint64_t groups,
                         u64 *__restrict__ out) {
    const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (g < groups) out[g] = *(const u64 *)(mt + aj_offset(row, (g0 + g) * 8, RT));
}

MAM_API int mam_ajtai_i8_matrix_row(void *handle, int row, uint64_t col0, uint64_t count, int8_t *h_out) {
    AjtaiI8 *A = (AjtaiI8 *)handle;
    if (!A || row < 0 || row >= A->rows || col0 + count > A->n || (count && !h_out)) return -2;
    if (count == 0) return 0;
    hipStream_t s = mam_stream();
    const uint64_t g0 = col0 / 8, g1 = (col0 + count + 7) / 8, groups = g1 - g0;
    u64 *d = nullptr;
    if (hipMallocAsync((void **)&d, groups * sizeof(u64), s) != hipSuccess) return -1;
    std::vector<u64> h(groups);
    k_aj_row<<<grid_1d(groups), MAM_BLOCK, 0, s>>>((const uint8_t *)A->mt, row, A->rows / 16, g0, groups, d);
    int rc = hipGetLastError() == hipSuccess ? 0 : -1;
    if (!rc && hipMemcpyAsync(h.data(), d, groups * sizeof(u64), hipMemcpyDeviceToHost, s) != hipSuccess) rc = -1;
    hipFreeAsync(d, s);
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = -1;
    if (rc) return rc;
    memcpy(h_out, (const uint8_t *)h.data() + (col0 - g0 * 8), count);
    return 0;
}

MAM_API int mam_ajtai_i8_free(void *handle) {
    AjtaiI8 *A = (AjtaiI8 *)handle;
    if (!A) return 0;
    (void)hipStreamSynchronize(mam_stream());
    (void)hipFree(A->mt);
    (void)hipFree(A->rowsum);
    delete A;
    return 0;
}

// ---------------------------------------------------------------------------
// Tensor-core commit probe (almost-goldilocks-cuda-rs/src/ffi.rs:627-633,
// cuda_almost_goldilocks/ajtai_tc_commit_probe.cuh:38-99): int8 GEMM partials
//   partial[kb][nt][r * 16 + c] = sum_{k in chunk kb} z[r][k] * M[k][nt * 16 + c]
// z row-major [16][K_total], M column-major [K_total][960] (entry (k, n) at n * K_total + k),
// chunk kb = [1024 kb, min(1024 (kb + 1), K_total)), nt < 60; the reference's throughput
// gate for a tensor-core commit. One wave per 16 x 16 output tile and chunk, on
// v_mfma_i32_16x16x64_i8: lane l = i + 16 g loads the 16 bytes k0 + 16 g .. + 15 of row i
// of z (A) and of column i of M (B); the same (g, byte) -> k assignment on both sides makes
// the instruction's internal k order irrelevant. K_total must be a multiple of 16.
// ---------------------------------------------------------------------------
static constexpr int TC_OUT_TILES = 60;   // 960 outputs = 42 ring rows x 64 coefficients, rounded to 16
static constexpr int TC_CHUNK_K = 1024;

__global__ __launch_bounds__(64) void k_tc_probe(const int8_t *__restrict__ z, const int8_t *__restrict__ M,
                                                 int32_t *__restrict__ partial, long K) {
    const int nt = blockIdx.x, kb = blockIdx.y, lane = threadIdx.x;
    const int i = lane & 15, g = lane >> 4;
    const long k_lo = (long)kb * TC_CHUNK_K;
    long k_hi = k_lo + TC_CHUNK_K;
    if (k_hi > K) k_hi = K;
    const int8_t *zr = z + (long)i * K;
    const int8_t *mc = M + (long)(nt * 16 + i) * K;
    v4i acc = v4i{0, 0, 0, 0};
    for (long k0 = k_lo; k0 < k_hi; k0 += 64) {
        const long k = k0 + 16 * g;
        v4i a = v4i{0, 0, 0, 0}, b = v4i{0, 0, 0, 0};
        if (k < k_hi) {
            a = *(const v4i *)(zr + k);
            b = *(const v4i *)(mc + k);
        }
        acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, acc, 0, 0, 0);
    }
    int32_t *o = partial + ((long)kb * TC_OUT_TILES + nt) * 256;
#pragma unroll
    for (int v = 0; v < 4; v++) o[(4 * g + v) * 16 + i] = acc[v];
}

MAM_API int ajtai_tc_commit_probe_ffi(const int8_t *d_z_int8, const int8_t *d_M_int8, int32_t *d_partial, int K_total,
                                      int num_K_chunks) {
    if (!d_z_int8 || !d_M_int8 || !d_partial || K_total <= 0 || K_total % 16 || num_K_chunks <= 0) return -1;
    k_tc_probe<<<dim3(TC_OUT_TILES, num_K_chunks), 64, 0, mam_stream()>>>(d_z_int8, d_M_int8, d_partial, K_total);
    return mam_launch_status();
}
