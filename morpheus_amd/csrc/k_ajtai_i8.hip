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
// * M is expanded once (setup) into HBM pre-tiled in MFMA A-fragment order: for every
//   32-column k-step and 16-row tile, the 64 lanes' 8-byte fragments are 512 contiguous
//   bytes, so each wave streams M with fully coalesced 8-byte loads. Row sums are
//   accumulated during expansion.
// * The witnesses become 8*B limb planes of n bytes (B-fragments: 8 contiguous bytes).
// * The GEMM runs v_mfma_i32_16x16x32_i8 over K-chunks of 2^14 columns (|chunk sum| <=
//   2^14 * 2^14 = 2^28, exact in int32); one workgroup per chunk, one wave per row tile,
//   one accumulator tile per 16 limb columns; int32 chunk partials are then summed in int64
//   and recombined mod p. At B = 1 the kernel streams the 1 GiB matrix once (HBM-bound);
//   the A fragment is reused across all column tiles, so larger B moves it toward MFMA.
#include <string.h>

#include "chacha.h"
#include "common.h"
#include "profile.h"

typedef int v4i __attribute__((ext_vector_type(4)));

static constexpr int AJ_TILE_BYTES = 512;        // 16 rows x 32 columns
static constexpr size_t AJ_KCHUNK = (size_t)1 << 14;  // columns per workgroup
static constexpr int AJ_MAX_BATCH = 32;

// byte offset of M[i][j] in the tiled layout with RT row tiles
__host__ __device__ __forceinline__ size_t aj_offset(size_t i, size_t j, int RT) {
    const size_t kt = j >> 5, kc = j & 31, rt = i >> 4, r = i & 15;
    return (kt * RT + rt) * AJ_TILE_BYTES + (((kc >> 3) << 4) + r) * 8 + (kc & 7);
}

// One thread per (row, ChaCha block of 64 columns).
__global__ void k_aj_expand(const u32 *__restrict__ key_words, int rows, size_t n, u64 *__restrict__ mt,
                            unsigned long long *__restrict__ rowsum) {
    const size_t blocks_per_row = n / 64;
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= (size_t)rows * blocks_per_row) return;
    const u32 i = (u32)(t / blocks_per_row);
    const u32 b = (u32)(t % blocks_per_row);
    u32 key[8];
#pragma unroll
    for (int k = 0; k < 8; k++) key[k] = key_words[k];
    const u32 nonce[3] = {i, 0, 0};
    u32 w[16];
    chacha::block<8>(key, b, nonce, w);
    const int RT = rows / 16;
    long long sum = 0;
#pragma unroll
    for (int g = 0; g < 8; g++) {
        const u64 v = ((u64)w[2 * g + 1] << 32) | w[2 * g];  // columns 64b + 8g .. + 7
        mt[aj_offset(i, (size_t)b * 64 + 8 * g, RT) / 8] = v;
#pragma unroll
        for (int k = 0; k < 8; k++) sum += (signed char)(v >> (8 * k));
    }
    atomicAdd(&rowsum[i], (unsigned long long)sum);
}

// planes[(w * 8 + l) * n + j] = byte l of canonical z_w[j], XOR 0x80 (= byte - 128 as int8).
// One thread per (witness, 8 consecutive columns): reads 64 bytes, writes one word per plane.
__global__ void k_aj_limbs(const u64 *__restrict__ z, int batch, size_t n, u64 *__restrict__ planes) {
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t groups = n / 8;
    if (t >= (size_t)batch * groups) return;
    const size_t w = t / groups, g = t % groups;
    u64 v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = gl_canon(z[w * n + 8 * g + k]);
#pragma unroll
    for (int l = 0; l < 8; l++) {
        u64 out = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) out |= ((v[k] >> (8 * l)) & 0xFF) << (8 * k);
        planes[((w * 8 + l) * n) / 8 + g] = out ^ 0x8080808080808080ULL;
    }
}

// grid: one workgroup per K-chunk; block: 64 * RT threads (wave w = row tile w).
// partial[(chunk * rows + row) * (CT * 16) + col] = sum over the chunk (int32).
template <int CT>
__global__ __launch_bounds__(1024) void k_aj_gemm(const u64 *__restrict__ mt, const u64 *__restrict__ planes,
                                                  int RT, size_t n, int nplanes, int *__restrict__ partial) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const size_t steps = n >> 5;
    const size_t kt0 = (size_t)blockIdx.x * (AJ_KCHUNK >> 5);
    size_t kt1 = kt0 + (AJ_KCHUNK >> 5);
    if (kt1 > steps) kt1 = steps;
    v4i acc[CT];
#pragma unroll
    for (int c = 0; c < CT; c++) acc[c] = v4i{0, 0, 0, 0};
    const int q = lane >> 4;  // k sub-block of this lane's fragment
    const int lc = lane & 15;
    for (size_t kt = kt0; kt < kt1; kt++) {
        const u64 a = mt[(kt * RT + w) * (AJ_TILE_BYTES / 8) + lane];
#pragma unroll
        for (int c = 0; c < CT; c++) {
            const int col = c * 16 + lc;
            const u64 b = col < nplanes ? planes[((size_t)col * n + kt * 32 + q * 8) / 8] : 0;
            acc[c] = __builtin_amdgcn_mfma_i32_16x16x32_i8((long)a, (long)b, acc[c], 0, 0, 0);
        }
    }
    const int rows = RT * 16;
#pragma unroll
    for (int c = 0; c < CT; c++) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = w * 16 + q * 4 + v;
            const int col = c * 16 + lc;
            partial[((size_t)blockIdx.x * rows + row) * (CT * 16) + col] = acc[c][v];
        }
    }
}

// out[w * rows + i] = sum_l 2^(8l) (S'_{i,(w,l)} + 128 R_i) mod p; one thread per (w, i).
__global__ void k_aj_finish(const int *__restrict__ partial, const unsigned long long *__restrict__ rowsum, int rows,
                            int cols, int chunks, int batch, u64 *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= batch * rows) return;
    const int w = t / rows, i = t % rows;
    const long long R = (long long)rowsum[i];
    u64 acc = 0;
    for (int l = 0; l < 8; l++) {
        const int col = w * 8 + l;
        long long s = 0;
        for (int c = 0; c < chunks; c++) s += partial[((size_t)c * rows + i) * cols + col];
        s += 128 * R;  // |s| < 2^38 < p
        const u64 r = s >= 0 ? (u64)s : GL_P - (u64)(-s);
        acc = gl_add(acc, gl_mul(r, (u64)1 << (8 * l)));
    }
    out[(size_t)w * rows + i] = acc;
}

namespace {
struct AjtaiI8 {
    int rows = 0;
    size_t n = 0;
    u64 *mt = nullptr;                     // rows * n bytes, tiled
    unsigned long long *rowsum = nullptr;  // rows
    u64 *planes = nullptr;                 // planes_batch * 8 * n bytes, grown on demand
    int planes_batch = 0;
    int *partial = nullptr;                // chunks * rows * 16 * ceil(batch / 2)
};
}  // namespace

// Expand the matrix for (key, rows, n): rows a multiple of 16 (<= 256), n a multiple of 64.
MAM_API int mam_ajtai_i8_setup(const uint8_t *key32, int rows, uint64_t n, void **handle) {
    if (!handle || !key32 || rows <= 0 || rows % 16 || rows > 256 || n == 0 || n % 64) return -2;
    *handle = nullptr;
    AjtaiI8 *A = new AjtaiI8();
    A->rows = rows;
    A->n = n;
    const size_t chunks = (n + AJ_KCHUNK - 1) / AJ_KCHUNK;
    hipStream_t s = mam_stream();
    u32 *d_key = nullptr;
    int rc = 0;
    if (hipMalloc((void **)&A->mt, (size_t)rows * n) != hipSuccess ||
        hipMalloc((void **)&A->rowsum, rows * sizeof(u64)) != hipSuccess ||
        hipMalloc((void **)&A->partial, chunks * rows * 16 * (AJ_MAX_BATCH / 2) * sizeof(int)) != hipSuccess ||
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
        k_aj_expand<<<grid_1d(threads), MAM_BLOCK, 0, s>>>(d_key, rows, n, A->mt, A->rowsum);
        if (hipGetLastError() != hipSuccess) rc = -1;
    }
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = -1;
    if (d_key) (void)hipFree(d_key);
    if (rc) {
        if (A->mt) (void)hipFree(A->mt);
        if (A->rowsum) (void)hipFree(A->rowsum);
        if (A->partial) (void)hipFree(A->partial);
        delete A;
        return rc;
    }
    *handle = A;
    return 0;
}

// d_z: batch witnesses of n words back to back (device); d_out: batch * rows words (device).
MAM_API int mam_ajtai_i8_commit(void *handle, const u64 *d_z, int batch, u64 *d_out) {
    AjtaiI8 *A = (AjtaiI8 *)handle;
    if (!A || !d_z || !d_out || batch < 1 || batch > AJ_MAX_BATCH) return -2;
    hipStream_t s = mam_stream();
    const size_t n = A->n;
    const int RT = A->rows / 16;
    const size_t chunks = (n + AJ_KCHUNK - 1) / AJ_KCHUNK;
    const int nplanes = 8 * batch;
    int ct = (nplanes + 15) / 16;
    int CT = ct <= 1 ? 1 : ct <= 2 ? 2 : ct <= 4 ? 4 : ct <= 8 ? 8 : 16;
    if (batch > A->planes_batch) {
        if (A->planes) {
            (void)hipStreamSynchronize(s);
            (void)hipFree(A->planes);
            A->planes = nullptr;
            A->planes_batch = 0;
        }
        if (hipMalloc((void **)&A->planes, (size_t)batch * 8 * n) != hipSuccess) return -1;
        A->planes_batch = batch;
    }
    k_aj_limbs<<<grid_1d((size_t)batch * (n / 8)), MAM_BLOCK, 0, s>>>(d_z, batch, n, A->planes);
    if (hipGetLastError() != hipSuccess) return -1;
    const unsigned grid = (unsigned)chunks;
    const int threads = 64 * RT;
    ProfScope prof(PROF_AJTAI_GEMM, (double)A->rows * n + (double)nplanes * n, s);
    switch (CT) {
        case 1: k_aj_gemm<1><<<grid, threads, 0, s>>>(A->mt, A->planes, RT, n, nplanes, A->partial); break;
        case 2: k_aj_gemm<2><<<grid, threads, 0, s>>>(A->mt, A->planes, RT, n, nplanes, A->partial); break;
        case 4: k_aj_gemm<4><<<grid, threads, 0, s>>>(A->mt, A->planes, RT, n, nplanes, A->partial); break;
        case 8: k_aj_gemm<8><<<grid, threads, 0, s>>>(A->mt, A->planes, RT, n, nplanes, A->partial); break;
        default: k_aj_gemm<16><<<grid, threads, 0, s>>>(A->mt, A->planes, RT, n, nplanes, A->partial); break;
    }
    if (hipGetLastError() != hipSuccess) return -1;
    k_aj_finish<<<grid_1d((size_t)batch * A->rows), MAM_BLOCK, 0, s>>>(A->partial, A->rowsum, A->rows, CT * 16,
                                                                        (int)chunks, batch, d_out);
    return mam_sync_status();
}

// Copy `count` matrix entries of one row starting at column col0 to host (int8).
MAM_API int mam_ajtai_i8_matrix_row(void *handle, int row, uint64_t col0, uint64_t count, int8_t *h_out) {
    AjtaiI8 *A = (AjtaiI8 *)handle;
    if (!A || row < 0 || row >= A->rows || col0 + count > A->n) return -2;
    if (hipStreamSynchronize(mam_stream()) != hipSuccess) return -1;
    const int RT = A->rows / 16;
    for (uint64_t j = col0; j < col0 + count; j += 8) {
        u64 v;
        if (hipMemcpy(&v, (const uint8_t *)A->mt + aj_offset(row, j & ~7ull, RT), 8, hipMemcpyDeviceToHost) !=
            hipSuccess)
            return -1;
        for (uint64_t k = 0; k < 8; k++) {
            uint64_t jj = (j & ~7ull) + k;
            if (jj >= col0 && jj < col0 + count) h_out[jj - col0] = (int8_t)(v >> (8 * k));
        }
    }
    return 0;
}

MAM_API int mam_ajtai_i8_free(void *handle) {
    AjtaiI8 *A = (AjtaiI8 *)handle;
    if (!A) return 0;
    (void)hipStreamSynchronize(mam_stream());
    (void)hipFree(A->mt);
    (void)hipFree(A->rowsum);
    (void)hipFree(A->planes);
    (void)hipFree(A->partial);
    delete A;
    return 0;
}
