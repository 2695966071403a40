// Matmul proof C = A B over Goldilocks (the benchmark's fifth configuration), composed from
// the reference's einsum proof for "ij,jk->ik" (zk-torch-3/src/basicblock/einsum.rs:807-1098):
// the output claim C(r) at a random point reduces, by one sumcheck over the inner index,
// to A and B evaluated at points, and every matrix is committed with Basefold and opened at
// its point on the same Fiat-Shamir transcript.
//
// Layout: A is m x k, B is k x n, C is m x n, row-major, m = 2^lm, k = 2^lk, n = 2^ln;
// index bits are LSB first, so A's low lk bits are its column (the inner index l) and its
// high lm bits its row i; B's low ln bits are its column j; C's low ln bits are j.
//
//   T = Transcript::new(label)
//   for X in A, B, C: root_X = BasefoldCommitment::commit(X); T.append_scalars("matmul-root", root_X)
//   [optional] a = int8 Ajtai commitment of C (first |C| matrix columns); T.append_scalars("matmul-ajtai", a)
//   r = (r_j, r_i) = ln + lm draws of T.challenge_ext2("matmul-point")
//   claim = C(r);  T.append_ext2("matmul-claim", claim)
//   A'(l) = A(r_i, l)   fused permute + partial evaluation (the row variables moved first)
//   B'(l) = B(l, r_j)   partial evaluation of the low ln variables
//   sumcheck over l of [A', B', 1] (degree 3; the einsum's "ones" polynomial: matmul has no
//     index shared by both inputs and the output)  -> r_l, final evaluations
//   open A at (r_l, r_i), B at (r_j, r_l), C at (r_j, r_i), in that order, on T
//
// The verifier checks claim = sum of the first round message's g(0) + g(1), the sumcheck,
// and the three openings (whose evaluations are final_evals[0], final_evals[1], claim).
#include <string.h>

#include <vector>

#include "common.h"
#include "matmul.h"
#include "transcript.h"

extern "C" {
int mam_sumcheck_prove(mam_transcript *t, const u64 *d_polys, int num_polys, int num_vars, int input_is_base,
                       u64 *round_msgs, u64 *challenges, u64 *final_evals, u64 *final_eval);
int mam_basefold_commit_into(const u64 *d_evals, int num_vars, int log_rate, int hash, u64 *d_codeword,
                             u64 *d_bh_evals, u64 *d_tree, u64 *root_out);
int mam_basefold_commitment_wrap(u64 *d_codeword, u64 *d_bh_evals, u64 *d_tree, const u64 *root4, int num_vars,
                                 int log_rate, int hash, void **handle);
int mam_basefold_commitment_free(void *handle);
int mam_basefold_open_ext2(void *handle, const u64 *point2, const u64 *d_table, int table_nv, mam_transcript *t,
                           int num_queries, u64 *proof);
int partial_eval_ext2_from_gl_ffi(const u64 *d_input, u64 *d_output, const u64 *d_r, int log_n, int m);
int fused_permute_partial_eval_ffi(const u64 *d_evals, u64 *d_output, const u64 *d_eq_table, const u32 *d_lo_lut,
                                   const u32 *d_hi_lut, int n, int m, int half, int output_size, int smem_bytes);
int mam_ajtai_i8_commit_prefix(void *handle, const u64 *d_z, int batch, uint64_t n_used, u64 *d_out);
}
int mam_eq_table(const u64 *d_r, int nv, u64 *d_out, bool ext, hipStream_t s);


namespace {

const uint8_t L_ROOT[] = "matmul-root";
const uint8_t L_AJTAI[] = "matmul-ajtai";
const uint8_t L_POINT[] = "matmul-point";
const uint8_t L_CLAIM[] = "matmul-claim";

// Device buffers of one prover thread, grown on demand.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    u64 *get(size_t words) {
        const size_t want = (words ? words : 1) * sizeof(u64);
        if (want > bytes) {
            (void)hipStreamSynchronize(mam_stream());
            if (p) (void)hipFree(p);
            p = nullptr;
            bytes = 0;
            if (hipMalloc(&p, want) != hipSuccess) return nullptr;
            bytes = want;
        }
        return (u64 *)p;
    }
};

struct MatmulBufs {
    DevBuf cw[3], bh[3], tree[3];
    DevBuf polys, fold, eq, small, luts;
};
thread_local MatmulBufs t_mm;

// Host-to-device copy that has completed (so the host buffer may be reused) on return.
int upload(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s) != hipSuccess) return -1;
    return hipStreamSynchronize(s) == hipSuccess ? 0 : -1;
}

// p[2i] = 1, p[2i + 1] = 0 for i < k (the all-ones Ext2 polynomial)
__global__ void k_ones_ext2(u64 *p, size_t k) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= k) return;
    p[2 * i] = 1;
    p[2 * i + 1] = 0;
}

}  // namespace

// Proves one instance (see the file comment). ajtai_handle: an int8 Ajtai set-up with at
// least 2^(lm+ln) columns and ajtai_rows rows, or null. Returns 0 or the first failing status.
MAM_API int mam_matmul_prove(mam_matmul_instance *I, const u64 *d_table, int table_nv, int log_rate, int hash,
                             int num_queries, void *ajtai_handle, int ajtai_rows) {
    if (!I || !d_table || I->lm < 1 || I->lk < 1 || I->ln < 1) return -2;
    if (log_rate < 0 || log_rate > 4 || (hash != 0 && hash != 1) || num_queries < 0) return -2;
    const int lm = I->lm, lk = I->lk, ln = I->ln;
    const int nvs[3] = {lm + lk, lk + ln, lm + ln};
    for (int x = 0; x < 3; x++)
        if (nvs[x] < 2 || nvs[x] > table_nv || nvs[x] > 30) return -2;
    hipStream_t s = mam_stream();
    MatmulBufs &B = t_mm;
    const u64 *mats[3] = {I->d_a, I->d_b, I->d_c};
    u64 *cw[3], *bh[3], *tree[3];
    for (int x = 0; x < 3; x++) {
        const size_t cwl = (size_t)1 << (nvs[x] + log_rate);
        cw[x] = B.cw[x].get(cwl);
        bh[x] = B.bh[x].get((size_t)1 << nvs[x]);
        tree[x] = B.tree[x].get((cwl - 1) * 4);
        if (!cw[x] || !bh[x] || !tree[x]) return -1;
    }
    const size_t k = (size_t)1 << lk;
    u64 *polys = B.polys.get(3 * 2 * k);         // [A' | B' | ones], Ext2
    // partial-evaluation output: its first halving writes 2^(nv - 1) Ext2 = 2^nv words
    u64 *fold = B.fold.get((size_t)1 << (nvs[1] > nvs[2] ? nvs[1] : nvs[2]));
    u64 *eq = B.eq.get(2 * ((size_t)1 << lm));  // eq(r_i, .)
    u64 *small = B.small.get(2 * (lm + ln) + 2 * ln + 2 + 256);
    const int nA = lm + lk, half = nA / 2;
    u32 *luts = (u32 *)B.luts.get((((size_t)1 << half) + ((size_t)1 << (nA - half)) + 1) / 2 + 1);
    if (!polys || !fold || !eq || !small || !luts) return -1;
    u64 *d_point = small;                      // r = (r_j, r_i)
    u64 *d_claim = small + 2 * (lm + ln);      // 2 words
    u64 *d_ajtai = d_claim + 2;                // up to 256 words

    mam_transcript t;
    mam::tr_init(&t, I->label, I->label_len);
    int rc = 0;
    for (int x = 0; x < 3 && !rc; x++) {
        rc = mam_basefold_commit_into(mats[x], nvs[x], log_rate, hash, cw[x], bh[x], tree[x], I->roots + 4 * x);
        if (!rc) mam::tr_append_scalars(&t, L_ROOT, sizeof(L_ROOT) - 1, I->roots + 4 * x, 4);
    }
    if (!rc && ajtai_handle && I->ajtai) {
        if (ajtai_rows < 1 || ajtai_rows > 256) return -2;
        rc = mam_ajtai_i8_commit_prefix(ajtai_handle, I->d_c, 1, (uint64_t)1 << nvs[2], d_ajtai);
        if (!rc) rc = (hipMemcpyAsync(I->ajtai, d_ajtai, ajtai_rows * sizeof(u64), hipMemcpyDeviceToHost, s) ==
                           hipSuccess && hipStreamSynchronize(s) == hipSuccess) ? 0 : -1;
        if (!rc) mam::tr_append_scalars(&t, L_AJTAI, sizeof(L_AJTAI) - 1, I->ajtai, ajtai_rows);
    }
    if (rc) return rc;
    for (int v = 0; v < ln + lm; v++) mam::tr_challenge_ext2(&t, L_POINT, sizeof(L_POINT) - 1, I->point + 2 * v);
    // claim = C(r)
    rc = upload(d_point, I->point, 2 * (ln + lm) * sizeof(u64), s);
    if (!rc) rc = partial_eval_ext2_from_gl_ffi(I->d_c, fold, d_point, nvs[2], nvs[2]);
    if (!rc) rc = mam_publish_wait(fold, 2, nullptr, 0, I->claim, s);
    if (rc) return rc;
    mam::tr_append_ext2(&t, L_CLAIM, sizeof(L_CLAIM) - 1, I->claim);
    // A'(l) = A(r_i, l): move A's lm row variables first, fold them against eq(r_i, .)
    {
        std::vector<u32> lut(((size_t)1 << half) + ((size_t)1 << (nA - half)));
        auto old_bit = [&](int p) { return p < lm ? lk + p : p - lm; };
        for (size_t x = 0; x < ((size_t)1 << half); x++) {
            u32 v = 0;
            for (int p = 0; p < half; p++) v |= (u32)((x >> p) & 1) << old_bit(p);
            lut[x] = v;
        }
        for (size_t y = 0; y < ((size_t)1 << (nA - half)); y++) {
            u32 v = 0;
            for (int p = 0; p < nA - half; p++) v |= (u32)((y >> p) & 1) << old_bit(half + p);
            lut[((size_t)1 << half) + y] = v;
        }
        rc = upload(luts, lut.data(), lut.size() * sizeof(u32), s);
        if (!rc) rc = mam_eq_table(d_point + 2 * ln, lm, eq, true, s);
        if (!rc)
            rc = fused_permute_partial_eval_ffi(I->d_a, polys, eq, luts, luts + ((size_t)1 << half), nA, lm, half,
                                                (int)k, 0);
    }
    // B'(l) = B(l, r_j) and the all-ones polynomial
    if (!rc) rc = partial_eval_ext2_from_gl_ffi(I->d_b, fold, d_point, nvs[1], ln);
    if (!rc)
        rc = hipMemcpyAsync(polys + 2 * k, fold, 2 * k * sizeof(u64), hipMemcpyDeviceToDevice, s) == hipSuccess ? 0
                                                                                                               : -1;
    if (!rc) {
        k_ones_ext2<<<grid_1d(k), MAM_BLOCK, 0, s>>>(polys + 4 * k, k);
        rc = mam_launch_status();
    }
    u64 fe[2];
    if (!rc) rc = mam_sumcheck_prove(&t, polys, 3, lk, 0, I->round_msgs, I->challenges, I->final_evals, fe);
    if (rc) return rc;
    // openings: A at (r_l, r_i), B at (r_j, r_l), C at (r_j, r_i)
    std::vector<u64> pa(2 * nvs[0]), pb(2 * nvs[1]);
    memcpy(pa.data(), I->challenges, 2 * lk * sizeof(u64));
    memcpy(pa.data() + 2 * lk, I->point + 2 * ln, 2 * lm * sizeof(u64));
    memcpy(pb.data(), I->point, 2 * ln * sizeof(u64));
    memcpy(pb.data() + 2 * ln, I->challenges, 2 * lk * sizeof(u64));
    const u64 *pts[3] = {pa.data(), pb.data(), I->point};
    u64 *proofs[3] = {I->proof_a, I->proof_b, I->proof_c};
    for (int x = 0; x < 3 && !rc; x++) {
        void *h = nullptr;
        rc = mam_basefold_commitment_wrap(cw[x], bh[x], tree[x], I->roots + 4 * x, nvs[x], log_rate, hash, &h);
        if (!rc) rc = mam_basefold_open_ext2(h, pts[x], d_table, table_nv, &t, num_queries, proofs[x]);
        if (h) mam_basefold_commitment_free(h);
    }
    return rc;
}
