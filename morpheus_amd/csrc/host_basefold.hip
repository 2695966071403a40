// Basefold polynomial commitment over Goldilocks with an Ext2 opening: folding table,
// commit, open_ext2, verify_ext2.
//
// Mirrors goldilocks-cuda-rs/src/basefold.rs:
//   BasefoldTable::generate   :281-321 (xorshift 13/7/17 draws reduced mod p; per level
//                             2^(nv+lr-i-1) entries (x0, w = 1/(x1 - x0)), x1 redrawn while
//                             equal to x0)
//   commit / commit_device    :382-427
//   open_ext2                 :578-863 (transcript order: root, point, oracle 0,
//                             then per round: challenge, next oracle, fold, root)
//   extract_ext2_queries      :2310-2402 (index = sample_challenge() % (cw_len / 2))
//   verify_ext2               :1759-1925
// The CPU twin zk-torch-3/src/commit/cpu_basefold.rs:235-425 produces the same proof.
//
// Flat proof layout (u64, canonical):
//   eval[2] | oracles[nv][3][2] | folded_roots[nv-1][4] | final_cw[2^lr][2] |
//   num_queries x { index | values[nv][2][2] | initial path[nv+lr-1][4] |
//                   folded path i = 0..nv-2: [nv+lr-2-i][4] }
//
// Device pipeline of the opening: eq table (one pass), bit reversal of bh_evals, round-0
// oracle, then per round one fused fold+oracle launch for (eq, bh) and one fused
// fold+Merkle pass for the codeword; the host reads back 10 words per round (oracle and
// root) to run the transcript. Query values and paths are gathered by one kernel.
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "basefold.h"
#include "common.h"
#include "merkle.h"
#include "monolith.h"
#include "poseidon2.h"
#include "transcript.h"

namespace {

long table_len(int nv, int lr, int nrounds) {
    long total = 0;
    for (int i = 0; i < nrounds; i++) total += 1L << (nv + lr - i - 1);
    return total;
}

long table_level_offset(int table_nv, int lr, int level) {
    long off = 0;
    for (int i = 0; i < level; i++) off += 1L << (table_nv + lr - i - 1);
    return off;
}

long proof_len(int nv, int lr, int nq) {
    long q = 1 + 4L * nv + 4L * (nv + lr - 1);
    for (int i = 0; i <= nv - 2; i++) q += 4L * (nv + lr - 2 - i);
    return 2 + 6L * nv + 4L * (nv - 1) + 2L * (1L << lr) + (long)nq * q;
}

long query_len(int nv, int lr) { return (proof_len(nv, lr, 1) - proof_len(nv, lr, 0)); }

struct Commitment {
    int nv = 0, lr = 0, hash = 0;
    u64 *codeword = nullptr;  // 2^(nv+lr) base words (Type1 order)
    u64 *bh_evals = nullptr;  // 2^nv base words (bit-reversed evaluations)
    u64 *tree = nullptr;      // (2L - 1) * 4 words, L = 2^(nv+lr-1)
    u64 root[4] = {0, 0, 0, 0};
    bool owns = false;
};

const int MAX_LEVELS = 40;

struct QuerySrc {
    const u64 *cw;                     // initial base codeword
    const u64 *tree0;                  // initial tree
    const u64 *fcw[MAX_LEVELS];        // folded codewords 0..nv-2 (Ext2)
    const u64 *ftree[MAX_LEVELS];      // their trees
    int nv, lr;
};

// One workgroup per query: gathers values and authentication paths into the flat layout.
__global__ void k_gather_queries(QuerySrc q, const u64 *__restrict__ idx, u64 *__restrict__ out, long qlen) {
    const int qi = blockIdx.x;
    const u64 leaf = idx[qi];
    u64 *o = out + (size_t)qi * qlen;
    const int nv = q.nv, lr = q.lr;
    if (threadIdx.x == 0) o[0] = leaf;
    u64 *vals = o + 1;
    // values: round 0 from the base codeword, rounds 1..nv-1 from folded codewords
    for (int r = threadIdx.x; r < nv; r += blockDim.x) {
        u64 *v = vals + 4 * r;
        if (r == 0) {
            v[0] = q.cw[2 * leaf];
            v[1] = 0;
            v[2] = q.cw[2 * leaf + 1];
            v[3] = 0;
        } else {
            const u64 pi = leaf >> r;
            const u64 *f = q.fcw[r - 1];
            v[0] = f[4 * pi];
            v[1] = f[4 * pi + 1];
            v[2] = f[4 * pi + 2];
            v[3] = f[4 * pi + 3];
        }
    }
    // paths: tree k (k = 0 initial, k >= 1 folded tree k-1) at leaf index leaf >> k
    u64 *p = vals + 4 * nv;
    for (int k = 0; k < nv; k++) {
        const int depth = nv + lr - 1 - k;  // levels below the root
        const u64 *tr = k == 0 ? q.tree0 : q.ftree[k - 1];
        const u64 leaves = (u64)1 << depth;
        for (int l = threadIdx.x; l < depth; l += blockDim.x) {
            // level l starts at digest 2*leaves - 2*(leaves >> l)
            const u64 lvl_off = 2 * leaves - 2 * (leaves >> l);
            const u64 node = ((leaf >> k) >> l) ^ 1;
            const u64 *d = tr + 4 * (lvl_off + node);
            u64 *w = p + 4 * l;
            w[0] = d[0];
            w[1] = d[1];
            w[2] = d[2];
            w[3] = d[3];
        }
        p += 4 * depth;
    }
}

u64 *pinned_scratch() {
    static thread_local u64 *buf = nullptr;
    if (!buf && hipHostMalloc((void **)&buf, 64 * sizeof(u64), hipHostMallocDefault) != hipSuccess) buf = nullptr;
    return buf;
}

int read_back(u64 *host, const u64 *dev, size_t words, hipStream_t s) {
    if (words <= (size_t)MAM_MAILBOX_WORDS) return mam_publish_wait(dev, (int)words, nullptr, 0, host, s);
    if (hipMemcpyAsync(host, dev, words * sizeof(u64), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
    return hipStreamSynchronize(s) == hipSuccess ? 0 : -1;
}

void observe_hash(mam_transcript *t, const u64 *h4) {
    for (int i = 0; i < 4; i++) mam::tr_append_scalars(t, nullptr, 0, h4 + i, 1);
}
void observe_ext2(mam_transcript *t, const u64 *v2) { mam::tr_append_ext2(t, nullptr, 0, v2); }
ext2 sample_ext2(mam_transcript *t) {
    u64 a[2];
    mam::tr_challenge_ext2(t, nullptr, 0, a);
    return ext2{a[0], a[1]};
}

int commit_into(Commitment &C, const u64 *d_evals, hipStream_t s) {
    if (mam_commit_codeword(d_evals, C.bh_evals, C.codeword, C.nv, C.lr, s)) return -1;
    const size_t leaves = (size_t)1 << (C.nv + C.lr - 1);
    int rc = mam_merkle_build(C.codeword, C.tree, leaves, C.hash, 0, s);
    if (rc) return rc;
    u64 *h = pinned_scratch();
    if (!h) return -1;
    if (read_back(h, C.tree + mam_merkle_root_offset(leaves), 4, s)) return -1;
    memcpy(C.root, h, sizeof(C.root));
    return 0;
}

void free_async(void *p, hipStream_t s) {
    if (p) hipFreeAsync(p, s);
}

// Keep freed stream-ordered allocations cached in the pool across calls.
void pool_setup() {
    static bool done = false;
    if (done) return;
    int dev = 0;
    hipGetDevice(&dev);
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t thr = ~0ull;
        hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
    done = true;
}

int open_ext2(const Commitment &C, const u64 *point2, const u64 *d_table, int table_nv, mam_transcript *t, int nq,
              u64 *proof) {
    const int nv = C.nv, lr = C.lr;
    if (nv < 2 || table_nv < nv) return -2;
    hipStream_t s = mam_stream();
    pool_setup();
    const size_t n = (size_t)1 << nv, cw_len = (size_t)1 << (nv + lr);
    u64 *pc = proof;
    // transcript preamble
    observe_hash(t, C.root);
    for (int i = 0; i < nv; i++) observe_ext2(t, point2 + 2 * i);

    // all device scratch of the opening is carved from the thread's arena (slot 0)
    const long qlen = query_len(nv, lr);
    size_t need = 0;
    auto words = [](size_t w) { return (w + 31) & ~(size_t)31; };
    {
        need += words(2 * nv) + words(2 * n) + words(n) + 2 * (words(n) + words(n / 2)) + words(RED_BLOCKS * 6 + 6) +
                words(8) + words(nq) + words((size_t)nq * qlen);
        size_t flen = cw_len / 2;
        for (int r = 0; r < nv; r++) {
            need += words(2 * flen);
            if (r < nv - 1) need += words((2 * (flen / 2) - 1) * 4);
            flen /= 2;
        }
    }
    u64 *arena = mam_dev_arena(need, 0);
    size_t used = 0;
    auto alloc = [&](size_t w) -> u64 * {
        if (!arena) return nullptr;
        u64 *p = arena + used;
        used += words(w);
        return p;
    };
    int rc = 0;
    u64 *d_point = alloc(2 * nv);
    u64 *eq = alloc(2 * n);
    u64 *bh = alloc(n);
    u64 *eqp[2] = {alloc(n), alloc(n / 2)};   // Ext2 ping-pong: n/2 then n/4 elements
    u64 *bhp[2] = {alloc(n), alloc(n / 2)};
    u64 *red = alloc(RED_BLOCKS * 6 + 6);
    u64 *cnt = alloc(8);
    std::vector<u64 *> fcw(nv), ftree(nv - 1);
    {
        size_t flen = cw_len / 2;
        for (int r = 0; r < nv; r++) {
            fcw[r] = alloc(2 * flen);
            if (r < nv - 1) ftree[r] = alloc((2 * (flen / 2) - 1) * 4);
            flen /= 2;
        }
    }
    if (!arena) rc = -1;
    u64 *h = pinned_scratch();
    if (!h) rc = -1;
    RedScratch rs = {red, (u32 *)cnt, red + RED_BLOCKS * 6};
    if (!rc && hipMemsetAsync(cnt, 0, 64, s) != hipSuccess) rc = -1;
    if (!rc && hipMemcpyAsync(d_point, point2, 2 * nv * sizeof(u64), hipMemcpyHostToDevice, s) != hipSuccess) rc = -1;
    if (!rc) rc = mam_eq_table(d_point, nv, eq, true, s);
    if (!rc) rc = mam_bitrev(C.bh_evals, bh, nv, 1, s);
    if (!rc) rc = mam_open_oracle0(eq, bh, n / 2, rs, s);
    if (!rc) rc = read_back(h, rs.result, 6, s);
    u64 *o_eval = pc;
    u64 *o_oracles = pc + 2;
    u64 *o_roots = o_oracles + 6L * nv;
    u64 *o_final = o_roots + 4L * (nv - 1);
    u64 *o_queries = o_final + 2L * (1L << lr);
    if (!rc) {
        memcpy(o_oracles, h, 6 * sizeof(u64));
        // eval = P(0) + P(1) = 2 c0 + c1 + c2
        ext2 c0{h[0], h[1]}, c1{h[2], h[3]}, c2{h[4], h[5]};
        ext2 ev = e2_add(e2_add(c0, c0), e2_add(c1, c2));
        o_eval[0] = ev.c0;
        o_eval[1] = ev.c1;
        for (int k = 0; k < 3; k++) observe_ext2(t, o_oracles + 2 * k);
    }
    // round 0: mixed transition
    const u64 *eq_cur = eq, *bh_cur = bh;
    bool bh_base = true;
    size_t sc_size = n;
    size_t flen = cw_len / 2;  // folded codeword length of this round
    const u64 *cw_in = C.codeword;
    bool cw_base = true;
    for (int round = 0; round < nv - 1 && !rc; round++) {
        ext2 ch = sample_ext2(t);
        const int pp = round & 1;
        rc = mam_open_fold_oracle(eq_cur, bh_cur, bh_base, ch, eqp[pp], bhp[pp], sc_size / 4, rs, s);
        if (rc) break;
        eq_cur = eqp[pp];
        bh_cur = bhp[pp];
        bh_base = false;
        sc_size /= 2;
        const u64 *lvl = d_table + 2 * table_level_offset(table_nv, lr, round);
        rc = mam_fold_and_tree(cw_in, cw_base, lvl, ch, fcw[round], ftree[round], flen, C.hash, s);
        if (rc) break;
        const size_t leaves = flen / 2;
        // oracle (3 Ext2) and the new root in one mailbox round trip
        if (mam_publish_wait(rs.result, 6, ftree[round] + mam_merkle_root_offset(leaves), 4, h, s)) {
            rc = -1;
            break;
        }
        u64 *o = o_oracles + 6L * (round + 1);
        memcpy(o, h, 6 * sizeof(u64));
        for (int k = 0; k < 3; k++) observe_ext2(t, o + 2 * k);
        memcpy(o_roots + 4L * round, h + 6, 4 * sizeof(u64));
        observe_hash(t, o_roots + 4L * round);
        cw_in = fcw[round];
        cw_base = false;
        flen /= 2;
    }
    if (!rc) {
        // last fold, no tree
        ext2 ch = sample_ext2(t);
        const u64 *lvl = d_table + 2 * table_level_offset(table_nv, lr, nv - 1);
        rc = mam_cw_fold(cw_in, lvl, ch, fcw[nv - 1], flen, false, s);
        if (!rc) rc = read_back(h, fcw[nv - 1], 2 * flen, s);
        if (!rc) memcpy(o_final, h, 2 * flen * sizeof(u64));
    }
    if (!rc && nq > 0) {
        std::vector<u64> idx(nq);
        for (int q = 0; q < nq; q++) idx[q] = mam::tr_challenge_scalar(t, nullptr, 0) % (u64)(cw_len / 2);
        u64 *d_idx = alloc(nq);
        u64 *d_out = alloc((size_t)nq * qlen);
        if (!d_idx || !d_out) rc = -1;
        QuerySrc qs;
        memset(&qs, 0, sizeof(qs));
        qs.cw = C.codeword;
        qs.tree0 = C.tree;
        for (int r = 0; r < nv - 1; r++) {
            qs.fcw[r] = fcw[r];
            qs.ftree[r] = ftree[r];
        }
        qs.nv = nv;
        qs.lr = lr;
        if (!rc && hipMemcpyAsync(d_idx, idx.data(), nq * sizeof(u64), hipMemcpyHostToDevice, s) != hipSuccess)
            rc = -1;
        if (!rc) {
            k_gather_queries<<<nq, 64, 0, s>>>(qs, d_idx, d_out, qlen);
            if (hipGetLastError() != hipSuccess) rc = -1;
        }
        if (!rc && hipMemcpyAsync(o_queries, d_out, (size_t)nq * qlen * sizeof(u64), hipMemcpyDeviceToHost, s) !=
                       hipSuccess)
            rc = -1;
        if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = -1;
    }
    return rc;
}

// ---------------------------------------------------------------------------
// Host verifier (BasefoldVerifier::verify_ext2, basefold.rs:1759-1925)
// ---------------------------------------------------------------------------
void hash_gl_leaf(int hash, u64 a, u64 b, u64 *o) {
    if (hash == MAM_HASH_MONOLITH) mono::hash_gl_leaf(a, b, o);
    else p2::hash_gl_leaf(a, b, o);
}
void hash_ext2_leaf(int hash, const u64 *ab, u64 *o) {
    if (hash == MAM_HASH_MONOLITH) mono::hash_ext2_leaf(ab, o);
    else p2::hash_ext2_leaf(ab, o);
}
void hash_node(int hash, const u64 *l, const u64 *r, u64 *o) {
    if (hash == MAM_HASH_MONOLITH) mono::compress(l, r, o);
    else p2::compress(l, r, o);
}

bool verify_path(int hash, const u64 *leaf4, const u64 *path, int len, u64 index, const u64 *root4) {
    u64 cur[4], nxt[4];
    memcpy(cur, leaf4, sizeof(cur));
    for (int l = 0; l < len; l++) {
        const u64 *sib = path + 4 * l;
        if (index & 1) hash_node(hash, sib, cur, nxt);
        else hash_node(hash, cur, sib, nxt);
        memcpy(cur, nxt, sizeof(cur));
        index >>= 1;
    }
    for (int i = 0; i < 4; i++)
        if (gl_canon(cur[i]) != gl_canon(root4[i])) return false;
    return true;
}

ext2 ld2(const u64 *p) { return ext2{gl_canon(p[0]), gl_canon(p[1])}; }

int verify_ext2(const u64 *root4, const u64 *point2, int nv, int lr, int hash, const u64 *proof, int nq,
                const u64 *table, int table_nv, mam_transcript *t) {
    if (nv < 2 || table_nv < nv) return 0;
    const u64 *pc = proof;
    ext2 eval = ld2(pc);
    pc += 2;
    const u64 *oracles = pc;
    pc += 6L * nv;
    const u64 *roots = pc;
    pc += 4L * (nv - 1);
    const u64 *final_cw = pc;
    pc += 2L * (1L << lr);
    observe_hash(t, root4);
    for (int i = 0; i < nv; i++) observe_ext2(t, point2 + 2 * i);
    ext2 c0 = ld2(oracles), c1 = ld2(oracles + 2), c2 = ld2(oracles + 4);
    if (!e2_eq(e2_add(c0, e2_add(c0, e2_add(c1, c2))), eval)) return 0;
    for (int k = 0; k < 3; k++) observe_ext2(t, oracles + 2 * k);
    std::vector<ext2> chs(nv);
    for (int round = 0; round < nv - 1; round++) {
        ext2 ch = sample_ext2(t);
        chs[round] = ch;
        const u64 *o = oracles + 6L * round, *on = oracles + 6L * (round + 1);
        ext2 val = e2_add(ld2(o), e2_add(e2_mul(ld2(o + 2), ch), e2_mul(ld2(o + 4), e2_mul(ch, ch))));
        ext2 n0 = ld2(on), n1 = ld2(on + 2), n2 = ld2(on + 4);
        if (!e2_eq(val, e2_add(n0, e2_add(n0, e2_add(n1, n2))))) return 0;
        for (int k = 0; k < 3; k++) observe_ext2(t, on + 2 * k);
        observe_hash(t, roots + 4L * round);
    }
    chs[nv - 1] = sample_ext2(t);
    for (int q = 0; q < nq; q++) {
        (void)mam::tr_challenge_scalar(t, nullptr, 0);
        const u64 index = pc[0];
        const u64 *vals = pc + 1;
        const u64 *paths = vals + 4L * nv;
        if (index >= ((u64)1 << (nv + lr - 1))) return 0;
        u64 leaf[4];
        hash_gl_leaf(hash, vals[0], vals[2], leaf);
        if (!verify_path(hash, leaf, paths, nv + lr - 1, index, root4)) return 0;
        const u64 *path = paths + 4L * (nv + lr - 1);
        u64 fidx = index;
        for (int r = 0; r < nv; r++) {
            const u64 *lvl = table + 2 * table_level_offset(table_nv, lr, r);
            ext2 v0 = ld2(vals + 4L * r), v1 = ld2(vals + 4L * r + 2);
            ext2 cx = e2_sub(chs[r], e2_base(gl_canon(lvl[2 * fidx])));
            ext2 fr = e2_add(v0, e2_mul(e2_mul(cx, e2_sub(v1, v0)), e2_base(gl_canon(lvl[2 * fidx + 1]))));
            if (r < nv - 1) {
                ext2 nx0 = ld2(vals + 4L * (r + 1)), nx1 = ld2(vals + 4L * (r + 1) + 2);
                if (!e2_eq(fr, (fidx & 1) ? nx1 : nx0)) return 0;
                u64 ab[4] = {nx0.c0, nx0.c1, nx1.c0, nx1.c1}, lh[4];
                hash_ext2_leaf(hash, ab, lh);
                const int plen = nv + lr - 2 - r;
                if (!verify_path(hash, lh, path, plen, fidx / 2, roots + 4L * r)) return 0;
                path += 4L * plen;
            } else {
                if (!e2_eq(fr, ld2(final_cw + 2 * fidx))) return 0;
            }
            fidx /= 2;
        }
        pc = path;
    }
    return 1;
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
MAM_API long mam_basefold_table_len(int num_vars, int log_rate, int num_rounds) {
    return table_len(num_vars, log_rate, num_rounds);
}

MAM_API long mam_basefold_proof_len(int num_vars, int log_rate, int num_queries) {
    return proof_len(num_vars, log_rate, num_queries);
}

// Host table generation into entries[2 * table_len] as (point, weight) pairs.
MAM_API int mam_basefold_table_generate(int num_vars, int log_rate, int num_rounds, u64 seed, u64 *entries) {
    if (num_vars < 1 || log_rate < 0 || num_rounds < 1 || num_rounds > num_vars + log_rate) return -2;
    u64 st = seed;
    auto draw = [&]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return st % GL_P;
    };
    long off = 0;
    std::vector<u64> pref;
    for (int i = 0; i < num_rounds; i++) {
        const long lsize = 1L << (num_vars + log_rate - i - 1);
        u64 *e = entries + 2 * off;
        for (long k = 0; k < lsize; k++) {
            u64 x0 = draw(), x1 = draw();
            while (x1 == x0) x1 = draw();
            e[2 * k] = x0;
            e[2 * k + 1] = gl_sub(x1, x0);
        }
        // Montgomery batch inversion: same values as one Fermat inverse per entry
        pref.resize(lsize);
        u64 acc = 1;
        for (long k = 0; k < lsize; k++) {
            pref[k] = acc;
            acc = gl_mul(acc, e[2 * k + 1]);
        }
        u64 inv = gl_inv(acc);
        for (long k = lsize - 1; k >= 0; k--) {
            u64 d = e[2 * k + 1];
            e[2 * k + 1] = gl_mul(inv, pref[k]);
            inv = gl_mul(inv, d);
        }
        off += lsize;
    }
    return 0;
}

// Commit into caller buffers (BasefoldCommitment::commit_device without ownership):
// d_codeword 2^(nv+lr) words, d_bh_evals 2^nv words, d_tree (2^(nv+lr) - 1) * 4 words.
MAM_API int mam_basefold_commit_into(const u64 *d_evals, int num_vars, int log_rate, int hash, u64 *d_codeword,
                                     u64 *d_bh_evals, u64 *d_tree, u64 *root_out) {
    if (num_vars < 1 || log_rate < 0 || num_vars + log_rate > 34 || (hash != 0 && hash != 1)) return -2;
    Commitment C;
    C.nv = num_vars;
    C.lr = log_rate;
    C.hash = hash;
    C.codeword = d_codeword;
    C.bh_evals = d_bh_evals;
    C.tree = d_tree;
    int rc = commit_into(C, d_evals, mam_stream());
    if (!rc && root_out) memcpy(root_out, C.root, sizeof(C.root));
    return rc;
}

// Commit, allocating the commitment's device buffers (BasefoldCommitment::commit_device).
MAM_API int mam_basefold_commit(const u64 *d_evals, int num_vars, int log_rate, int hash, void **handle,
                                u64 *root_out) {
    if (!handle) return -2;
    *handle = nullptr;
    if (num_vars < 1 || log_rate < 0 || num_vars + log_rate > 34 || (hash != 0 && hash != 1)) return -2;
    Commitment *C = new Commitment();
    C->nv = num_vars;
    C->lr = log_rate;
    C->hash = hash;
    C->owns = true;
    const size_t cw = (size_t)1 << (num_vars + log_rate);
    if (hipMalloc((void **)&C->codeword, cw * sizeof(u64)) != hipSuccess ||
        hipMalloc((void **)&C->bh_evals, ((size_t)1 << num_vars) * sizeof(u64)) != hipSuccess ||
        hipMalloc((void **)&C->tree, (cw - 1) * 4 * sizeof(u64)) != hipSuccess) {
        hipFree(C->codeword);
        hipFree(C->bh_evals);
        hipFree(C->tree);
        delete C;
        return -1;
    }
    int rc = commit_into(*C, d_evals, mam_stream());
    if (rc) {
        hipFree(C->codeword);
        hipFree(C->bh_evals);
        hipFree(C->tree);
        delete C;
        return rc;
    }
    if (root_out) memcpy(root_out, C->root, sizeof(C->root));
    *handle = C;
    return 0;
}

// Wrap caller-owned commitment buffers (e.g. restored from a host cache) as a handle.
MAM_API int mam_basefold_commitment_wrap(u64 *d_codeword, u64 *d_bh_evals, u64 *d_tree, const u64 *root4,
                                         int num_vars, int log_rate, int hash, void **handle) {
    if (!handle || num_vars < 1 || log_rate < 0) return -2;
    Commitment *C = new Commitment();
    C->nv = num_vars;
    C->lr = log_rate;
    C->hash = hash;
    C->codeword = d_codeword;
    C->bh_evals = d_bh_evals;
    C->tree = d_tree;
    memcpy(C->root, root4, sizeof(C->root));
    C->owns = false;
    *handle = C;
    return 0;
}

MAM_API int mam_basefold_commitment_buffers(void *handle, u64 **d_codeword, u64 **d_bh_evals, u64 **d_tree,
                                            u64 *root_out) {
    Commitment *C = (Commitment *)handle;
    if (!C) return -2;
    if (d_codeword) *d_codeword = C->codeword;
    if (d_bh_evals) *d_bh_evals = C->bh_evals;
    if (d_tree) *d_tree = C->tree;
    if (root_out) memcpy(root_out, C->root, sizeof(C->root));
    return 0;
}

MAM_API int mam_basefold_commitment_free(void *handle) {
    Commitment *C = (Commitment *)handle;
    if (!C) return 0;
    if (C->owns) {
        hipFree(C->codeword);
        hipFree(C->bh_evals);
        hipFree(C->tree);
    }
    delete C;
    return 0;
}

// open_ext2: point2 (host) holds nv Ext2 values; d_table the device folding table built for
// table_nv >= nv with the commitment's log_rate; proof (host) has proof_len words.
MAM_API int mam_basefold_open_ext2(void *handle, const u64 *point2, const u64 *d_table, int table_nv,
                                   mam_transcript *t, int num_queries, u64 *proof) {
    Commitment *C = (Commitment *)handle;
    if (!C || !point2 || !d_table || !t || !proof || num_queries < 0) return -2;
    return open_ext2(*C, point2, d_table, table_nv, t, num_queries, proof);
}

// Returns 1 if the proof verifies, 0 otherwise (host only; table is the host copy).
MAM_API int mam_basefold_verify_ext2(const u64 *root4, const u64 *point2, int num_vars, int log_rate, int hash,
                                     const u64 *proof, int num_queries, const u64 *table, int table_nv,
                                     mam_transcript *t) {
    return verify_ext2(root4, point2, num_vars, log_rate, hash, proof, num_queries, table, table_nv, t);
}

// Host Monolith permutation (for checking the device one).
MAM_API void mam_monolith_permute_host(u64 *state12) { mono::permute(state12); }
