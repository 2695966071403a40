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
// Device pipeline of the opening (single point): the round-0 oracle from the stored
// bit-reversed bh, then per round the codeword fold and its Merkle tree with the round's
// bh fold and sums (product form) or eq / bh fold and oracle riding along the tree's
// launches; each round is queued behind a gate before the host has read the previous
// round's oracle and root, which come back through the pinned mailbox (10 words) for the
// transcript. Query values and paths are gathered by one kernel and copied back with the
// final codeword.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <vector>

#include "basefold.h"
#include "common.h"
#include "merkle.h"
#include "monolith.h"
#include "open_fold.h"
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
    // the codeword came from this library's commit transform, so it is 2^lr identical
    // blocks (buffers wrapped from elsewhere are not assumed to be)
    bool replicated = false;
};

const int MAX_LEVELS = 40;

struct QuerySrc {
    const u64 *cw;                     // initial codeword (base, or Ext2 when cw_ext2)
    int cw_ext2;
    const u64 *tree0;                  // initial tree
    const u64 *fcw[MAX_LEVELS];        // folded codewords 0..nv-2 (Ext2)
    const u64 *ftree[MAX_LEVELS];      // their trees
    int nv, lr;
    const u64 *fin;  // final codeword (fin_words), copied to fin_out by workgroup 0
    u64 *fin_out;
    int fin_words;
};

// Query indices passed in the kernel's arguments (no copy of their own), up to 128 queries.
static constexpr int QIDX_INLINE = 128;
struct QueryIdx {
    u64 v[QIDX_INLINE];
};

// One workgroup per query: gathers values and authentication paths into the flat layout.
// The indices come from `idx` in device memory, or from qi_inline when idx is null. Every
// 4-word item of the query (nv values, then the path entries of the nv trees, tree k at
// depth nv + lr - 1 - k) goes to its own thread, so all of a query's reads are in flight
// at once. `out` may be page-locked host memory (the writes then cross the bus directly).
static constexpr int GATHER_THREADS = 256;
__global__ __launch_bounds__(GATHER_THREADS) void k_gather_queries(QuerySrc q, const u64 *__restrict__ idx,
                                                                    u64 *__restrict__ out, long qlen,
                                                                    QueryIdx qi_inline) {
    const int qi = blockIdx.x;
    const u64 leaf = idx ? idx[qi] : qi_inline.v[qi];
    u64 *o = out + (size_t)qi * qlen;
    const int nv = q.nv, lr = q.lr;
    const int D = nv + lr - 1;                     // depth of the initial tree
    const int entries = nv * D - nv * (nv - 1) / 2;  // path entries of the nv trees
    if (threadIdx.x == 0) o[0] = leaf;
    if (qi == 0)
        for (int w = threadIdx.x; w < q.fin_words; w += blockDim.x) q.fin_out[w] = q.fin[w];
    u64 *vals = o + 1;
    for (int it = threadIdx.x; it < nv + entries; it += blockDim.x) {
        const u64 *d;
        u64 v[4];
        if (it < nv) {
            // value of round r: round 0 from the base codeword, rounds 1..nv-1 from folded codewords
            const int r = it;
            if (r == 0 && q.cw_ext2) {
                d = q.cw + 4 * leaf;
                v[0] = d[0], v[1] = d[1], v[2] = d[2], v[3] = d[3];
            } else if (r == 0) {
                v[0] = q.cw[2 * leaf], v[1] = 0, v[2] = q.cw[2 * leaf + 1], v[3] = 0;
            } else {
                d = q.fcw[r - 1] + 4 * (leaf >> r);
                v[0] = d[0], v[1] = d[1], v[2] = d[2], v[3] = d[3];
            }
        } else {
            // path entry e: tree k (k = 0 initial, k >= 1 folded tree k-1) at level l, leaf >> k
            int e = it - nv, k = 0;
            while (k < nv - 1 && e >= D - k) {
                e -= D - k;
                k++;
            }
            const int l = e, depth = D - k;
            const u64 *tr = k == 0 ? q.tree0 : q.ftree[k - 1];
            const u64 leaves = (u64)1 << depth;
            const u64 lvl_off = 2 * leaves - 2 * (leaves >> l);  // level l starts at this digest
            const u64 node = ((leaf >> k) >> l) ^ 1;
            d = tr + 4 * (lvl_off + node);
            v[0] = d[0], v[1] = d[1], v[2] = d[2], v[3] = d[3];
        }
        u64 *w = vals + 4 * (size_t)it;
        w[0] = v[0];
        w[1] = v[1];
        w[2] = v[2];
        w[3] = v[3];
    }
}

// 320 pinned words per thread: [0, 64) read-back words, [64, 192) host-to-device staging,
// [192, 320) the query indices of a pre-queued gather
static constexpr int PIN_IDX = 192;
u64 *pinned_scratch() {
    static thread_local u64 *buf = nullptr;
    if (!buf && hipHostMalloc((void **)&buf, 320 * sizeof(u64), hipHostMallocDefault) != hipSuccess) buf = nullptr;
    return buf;
}

// The opening's ticket counters (grid_finish_n): device words of the calling thread, zeroed
// once; every ticketed reduction leaves its counter at zero again, so no clear per opening.
u32 *ticket_counters() {
    static thread_local u32 *c = nullptr;
    if (!c) {
        u32 *p = nullptr;
        if (hipMalloc((void **)&p, 64) != hipSuccess) return nullptr;
        if (hipMemset(p, 0, 64) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
            (void)hipFree(p);
            return nullptr;
        }
        c = p;
    }
    return c;
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
    // the codeword is 2^lr copies of one block: hash one copy's subtree
    int rc = mam_merkle_build_copies(C.codeword, C.tree, leaves, C.hash, 0, 1 << C.lr, s);
    if (rc) return rc;
    C.replicated = true;
    u64 *h = pinned_scratch();
    if (!h) return -1;
    if (read_back(h, C.tree + mam_merkle_root_offset(leaves), 4, s)) return -1;
    memcpy(C.root, h, sizeof(C.root));
    return 0;
}

void free_async(void *p, hipStream_t s) {
    if (p) hipFreeAsync(p, s);
}

// The opening proper, from the eq table onwards: the sumcheck of eq * bh with its oracles,
// the folds of the codeword with their trees, the final codeword and the queries. Writes
// eval | oracles | folded roots | final cw | queries to `out`. open_ext2 runs it on one
// commitment (base bh and codeword); the batch opening runs it on the combined codeword
// (Ext2 bh and codeword), whose proof section has the same layout.
struct OpenSrc {
    const u64 *eq;     // Ext2 eq table at the opening point, natural order, 2^nv entries
    const u64 *bh;     // evaluations in natural order (base or Ext2)
    bool bh_base;
    const u64 *cw;     // codeword, Type1 order (base or Ext2), 2^(nv+lr) entries
    bool cw_base;
    const u64 *tree0;  // Merkle tree of cw
    int nv, lr, hash;
    // Product form (a single opening point, eq == nullptr): the point's Ext2 words on the
    // device and the host; rounds below K fold bh alone and return the sums U0, U1 of
    // mam_open_u, round K builds the scaled eq table the later (ride-along) rounds fold.
    const u64 *pt = nullptr, *pt_host = nullptr;
    int K = 0;
    // Product form only: bh (and every table the rounds fold) in bit-reversed order, as the
    // commitment stores it, with pt_rev the point's coordinates in reversed order on the
    // device (UJob::rev, OracleJob::rev); no bit reversal ahead of the rounds.
    bool rev = false;
    const u64 *pt_rev = nullptr;
    // Transcript preamble (root, then the point's nv Ext2 words) absorbed by open_core after
    // queuing the first launches, before it reads the first oracle (null: none)
    const u64 *pre_root = nullptr, *pre_point = nullptr;
};

// The round's oracle (c0, c1, c2) from U0, U1 when eq = sc * eq(r_k, r_k+1, ...) (product form).
static void oracle_from_u(const u64 *U, ext2 sc, ext2 rk, u64 *o6) {
    const ext2 u0{gl_canon(U[0]), gl_canon(U[1])}, u1{gl_canon(U[2]), gl_canon(U[3])};
    const ext2 a = e2_mul(sc, e2_sub(e2_one(), rk)), b = e2_mul(sc, rk);
    const ext2 c0 = e2_mul(a, u0), c2 = e2_mul(e2_sub(b, a), e2_sub(u1, u0));
    const ext2 c1 = e2_sub(e2_sub(e2_mul(b, u1), c0), c2);
    const ext2 c[3] = {c0, c1, c2};
    for (int k = 0; k < 3; k++) {
        o6[2 * k] = c[k].c0;
        o6[2 * k + 1] = c[k].c1;
    }
}

static bool open_rev() {
    static const bool v = [] {
        const char *e = getenv("MAM_OPEN_REV");
        return !(e && e[0] == '0');
    }();
    return v;
}

static bool uj_ride() {
    static const bool v = [] {
        const char *e = getenv("MAM_UJ_RIDE");
        return !(e && e[0] == '0');
    }();
    return v;
}

int open_core(const OpenSrc &S, const u64 *d_table, int table_nv, mam_transcript *t, int nq, u64 *out) {
    const int nv = S.nv, lr = S.lr;
    hipStream_t s = mam_stream();
    const size_t n = (size_t)1 << nv, cw_len = (size_t)1 << (nv + lr);
    // device scratch of the opening is carved from the thread's arena (slot 0)
    const long qlen = query_len(nv, lr);
    size_t need = 0;
    auto words = [](size_t w) { return (w + 31) & ~(size_t)31; };
    {
        need += 2 * (words(n) + words(n / 2)) + words(RED_BLOCKS * 6 + 6) + words(nq > 0 ? nq : 1) +
                words(2 * ((size_t)1 << lr) + (size_t)nq * qlen);
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
    u64 *eqp[2] = {alloc(n), alloc(n / 2)};   // Ext2 ping-pong: n/2 then n/4 elements
    u64 *bhp[2] = {alloc(n), alloc(n / 2)};
    u64 *red = alloc(RED_BLOCKS * 6 + 6);
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
    u32 *tickets = ticket_counters();
    if (!tickets) rc = -1;
    RedScratch rs = {red, tickets, red + RED_BLOCKS * 6};
    auto point_var = [&](int k) { return ext2{gl_canon(S.pt_host[2 * k]), gl_canon(S.pt_host[2 * k + 1])}; };
    // The rounds from K on ride along tree launches and fold an eq table. Their table is
    // scale_K * eq(r_K, ..., r_{nv-1}; .); the unscaled table depends on the point only and is
    // queued behind the first oracle's post, so the device builds it while the host runs its
    // first transcript step, and the host multiplies those rounds' oracles (linear in eq) by
    // scale_K.
    const bool late_eq = !S.eq && S.K < nv - 1;
    ext2 kscale = e2_one();
    // Pre-queued rounds (common.h gates): round r's launches, which read their challenge from
    // gate r's device slot, are queued before the host waits for round r - 1's results, and the
    // kernel that posts round r - 1's results waits for gate r before it ends, so the device
    // starts the round as soon as the host releases the challenge.
    const bool pq = mam_prequeue_enabled() && mam_gate_ready();
    const bool ride_all = uj_ride();
    struct Queued {
        u64 gate = 0, seq = 0;
        u64 *box = nullptr;
        GateJob g = {};
        const u64 *chp = nullptr;
    };
    std::vector<Queued> qd(nv);  // rounds 0 .. nv-2 and the final fold (nv - 1)
    // The device work of round r depends on r alone (and on the challenge, by value or at chp).
    // next: the gate the round's post waits for (null: none)
    auto queue_round = [&](int round, ext2 ch, const u64 *chp, const GateJob *next) -> int {
        const int pp = round & 1;
        const bool prod = !S.eq && round < S.K;
        const u64 *eq_in = round == 0 ? S.eq : eqp[pp ^ 1];  // round K: the late eq table
        const u64 *bh_in = round == 0 ? S.bh : bhp[pp ^ 1];
        const bool bhb = round == 0 && S.bh_base;
        const size_t scs = n >> round, fl = (cw_len / 2) >> round;
        const u64 *cwi = round == 0 ? S.cw : fcw[round - 1];
        const bool cwb = round == 0 && S.cw_base;
        const u64 *lvl = d_table + 2 * table_level_offset(table_nv, lr, round);
        if (round == nv - 1) return mam_cw_fold(cwi, lvl, ch, fcw[nv - 1], fl, cwb, s, chp);  // last fold, no tree
        if (prod && (nv - round - 2 > 30 || scs / 4 > ((size_t)1 << (nv - round - 2)))) return -2;
        // the eq / bh fold with the next oracle, run along with the folded codeword's tree
        const OracleJob job = {eq_in, bh_in, ch, eqp[pp], bhp[pp], scs / 4, rs.partials, rs.counter, rs.result, 0,
                               bhb ? 1 : 0, chp, S.rev ? 1 : 0};
        // product form: the bh fold and U sums ride along the tree's first launch (or run on a
        // launch of their own before it with MAM_UJ_RIDE=0)
        const u64 *ur = S.rev ? S.pt_rev : S.pt + 2 * (round + 2);  // E's coordinates
        const UJob ujob = {bh_in, bhp[pp], ch, scs / 4, ur, nv - round - 2, rs.partials,
                           rs.result, rs.counter, chp, S.rev ? 1 : 0};
        const bool ride = prod && ride_all;
        if (prod && !ride) {
            const int e = mam_open_u(bh_in, bhb, true, ch, bhp[pp], scs / 4, ur, nv - round - 2, rs, s, chp, S.rev);
            if (e) return e;
        }
        // the oracle (3 Ext2) and the new root come back in one mailbox round trip, posted by
        // the tree's final launch where it can
        Queued &q = qd[round];
        q.box = mam_mailbox_arm(&q.seq);
        if (!q.box) return -1;
        const int nres = prod ? 4 : 6;  // U0, U1 or the oracle
        const MailJob mail = {rs.result, nres, q.box, q.seq, next ? *next : GateJob{}};
        bool posted = false;
        const int e = mam_fold_and_tree(cwi, cwb, lvl, ch, fcw[round], ftree[round], fl, S.hash, s,
                                        prod ? nullptr : &job, &mail, &posted, ride ? &ujob : nullptr, chp);
        if (e) return e;
        if (!posted &&
            mam_mailbox_post(rs.result, nres, ftree[round] + mam_merkle_root_offset(fl / 2), 4, q.box, q.seq, s,
                             next))
            return -1;
        return 0;
    };
    auto arm = [&](int round) -> bool {
        qd[round].chp = mam_gate_arm(2, &qd[round].g, &qd[round].gate);
        return qd[round].chp != nullptr;
    };
    // pre-queue round r (behind gate r, armed before): a tree round's post carries gate r + 1
    auto queue_gated = [&](int round) -> int {
        if (round == nv - 1) return queue_round(round, e2_zero(), qd[round].chp, nullptr);
        if (!arm(round + 1)) return -1;
        return queue_round(round, e2_zero(), qd[round].chp, &qd[round + 1].g);
    };
    auto release = [&](int round, ext2 ch) {
        const u64 v[2] = {ch.c0, ch.c1};
        mam_gate_open(qd[round].gate, v, 2);
    };
    // The end of the proof: query gather (indices at idx, or in qin when idx is null) and one
    // copy of the final codeword and the queries into `out`.
    const size_t fin_words = 2 * ((size_t)1 << lr);
    u64 *d_out = alloc(fin_words + (size_t)nq * qlen);
    u64 *d_idx = alloc(nq > 0 ? nq : 1);
    if (!d_out || !d_idx) rc = -1;
    u64 *o_final_ptr = out + 2 + 6L * nv + 4L * (nv - 1);  // the queries follow the final codeword
    auto queue_copy = [&]() -> int {
        return hipMemcpyAsync(o_final_ptr, d_out, (fin_words + (size_t)nq * qlen) * sizeof(u64),
                              hipMemcpyDeviceToHost, s) == hipSuccess
                   ? 0
                   : -1;
    };
    auto queue_gather = [&](const u64 *idx, const QueryIdx &qin) -> int {
        if (nq > 0) {
            QuerySrc qs;
            memset(&qs, 0, sizeof(qs));
            qs.cw = S.cw;
            qs.cw_ext2 = S.cw_base ? 0 : 1;
            qs.tree0 = S.tree0;
            for (int r = 0; r < nv - 1; r++) {
                qs.fcw[r] = fcw[r];
                qs.ftree[r] = ftree[r];
            }
            qs.nv = nv;
            qs.lr = lr;
            qs.fin = fcw[nv - 1];
            qs.fin_out = d_out;
            qs.fin_words = (int)fin_words;
            k_gather_queries<<<nq, GATHER_THREADS, 0, s>>>(qs, idx, d_out + fin_words, qlen, qin);
            if (hipGetLastError() != hipSuccess) return -1;
        } else if (hipMemcpyAsync(d_out, fcw[nv - 1], fin_words * sizeof(u64), hipMemcpyDeviceToDevice, s) !=
                   hipSuccess) {
            return -1;
        }
        return 0;
    };
    // Pre-queued: the gather reads its indices from device memory, copied there by a gate
    // that the host releases once it has drawn them (at most 128 queries). The copy back is
    // queued with it only into page-locked memory: a copy into pageable memory waits on the
    // host for its data, which would stand behind the unreleased gate.
    u64 *pin_idx = h ? h + PIN_IDX : nullptr;
    const bool tail_pq = pq && nq > 0 && nq <= QIDX_INLINE && pin_idx;
    bool tail_queued = false, copy_queued = false;
    u64 idx_gate = 0;
    auto out_pinned = [&]() {
        hipPointerAttribute_t a;
        if (hipPointerGetAttributes(&a, o_final_ptr) != hipSuccess) {
            (void)hipGetLastError();  // pageable memory: not an error of this call's stream
            return false;
        }
        return a.type == hipMemoryTypeHost;
    };
    auto queue_tail = [&]() -> int {
        GateJob ig;
        if (!mam_gate_arm(1, &ig, &idx_gate)) return -1;
        ig.slot = pin_idx;
        ig.dst = d_idx;
        ig.n = nq;
        if (mam_gate_launch(ig, s)) return -1;
        tail_queued = true;
        int e = queue_gather(d_idx, QueryIdx{});
        if (!e && out_pinned()) {
            e = queue_copy();
            copy_queued = !e;
        }
        return e;
    };
    {
        u64 seq0 = 0;
        u64 *box0 = rc ? nullptr : mam_mailbox_arm(&seq0);
        if (!rc && !box0) rc = -1;
        const int n0 = S.eq ? 6 : 4;
        if (!rc)
            rc = S.eq ? mam_open_oracle0(S.eq, S.bh, S.bh_base, n / 2, rs, s)
                      : mam_open_u(S.bh, S.bh_base, false, e2_zero(), nullptr, n / 2, S.rev ? S.pt_rev : S.pt + 2,
                                   nv - 1, rs, s, nullptr, S.rev);
        if (!rc) rc = mam_mailbox_post(rs.result, n0, nullptr, 0, box0, seq0, s);
        if (!rc && late_eq) {
            if (S.rev) {
                // coordinates K .. nv-1 in reversed order: the table in bit-reversed order
                rc = mam_eq_table(S.pt_rev, nv - S.K, eqp[(S.K & 1) ^ 1], true, s);
            } else {
                const ext2 rk = point_var(S.K);
                rc = mam_eq_table_scaled(S.pt + 2 * S.K, nv - S.K, e2_sub(e2_one(), rk), rk, eqp[(S.K & 1) ^ 1], s);
            }
        }
        if (!rc && pq) rc = arm(0) ? mam_gate_launch(qd[0].g, s) : -1;
        if (!rc && pq) rc = queue_gated(0);
        if (S.pre_root) {
            observe_hash(t, S.pre_root);
            for (int i = 0; i < nv; i++) observe_ext2(t, S.pre_point + 2 * i);
        }
        if (!rc) rc = mam_mailbox_wait(box0, seq0, n0, h, s);
        if (!rc && !S.eq) oracle_from_u(h, e2_one(), point_var(0), h);
    }
    u64 *o_eval = out;
    u64 *o_oracles = out + 2;
    u64 *o_roots = o_oracles + 6L * nv;
    if (!rc) {
        memcpy(o_oracles, h, 6 * sizeof(u64));
        // eval = P(0) + P(1) = 2 c0 + c1 + c2
        ext2 c0{h[0], h[1]}, c1{h[2], h[3]}, c2{h[4], h[5]};
        ext2 ev = e2_add(e2_add(c0, c0), e2_add(c1, c2));
        o_eval[0] = ev.c0;
        o_eval[1] = ev.c1;
        for (int k = 0; k < 3; k++) observe_ext2(t, o_oracles + 2 * k);
    }
    ext2 scale = e2_one();  // product form: eq of this round = scale * eq(r_round, ...)
    for (int round = 0; round < nv - 1 && !rc; round++) {
        const ext2 ch = sample_ext2(t);
        const bool prod = !S.eq && round < S.K;
        ext2 next_scale = scale;
        if (!S.eq) {
            const ext2 rr = point_var(round);
            next_scale = e2_mul(scale, e2_add(e2_mul(e2_sub(e2_one(), rr), e2_sub(e2_one(), ch)), e2_mul(rr, ch)));
            // the rounds from K on ride along tree launches on the eq table built above
            if (round == S.K) kscale = scale;
        }
        if (pq) {
            release(round, ch);
            rc = queue_gated(round + 1);  // the next round (or the final fold) waits behind its gate
            if (!rc && round + 1 == nv - 1 && tail_pq) rc = queue_tail();
        } else {
            rc = queue_round(round, ch, nullptr, nullptr);
        }
        if (rc) break;
        const int nres = prod ? 4 : 6;
        if (mam_mailbox_wait(qd[round].box, qd[round].seq, nres + 4, h, s)) {
            rc = -1;
            break;
        }
        u64 *o = o_oracles + 6L * (round + 1);
        if (prod) {
            oracle_from_u(h, next_scale, point_var(round + 1), o);
        } else if (late_eq) {
            for (int k = 0; k < 3; k++) {
                const ext2 c = e2_mul(kscale, ext2{gl_canon(h[2 * k]), gl_canon(h[2 * k + 1])});
                o[2 * k] = c.c0;
                o[2 * k + 1] = c.c1;
            }
        } else {
            memcpy(o, h, 6 * sizeof(u64));
        }
        for (int k = 0; k < 3; k++) observe_ext2(t, o + 2 * k);
        memcpy(o_roots + 4L * round, h + nres, 4 * sizeof(u64));
        scale = next_scale;
        observe_hash(t, o_roots + 4L * round);
    }
    if (!rc) {
        // last fold, no tree; the query indices follow its challenge directly (the final
        // codeword is not observed), so the gather follows the fold and one copy brings back
        // the final codeword and the queries
        const ext2 ch = sample_ext2(t);
        if (pq) release(nv - 1, ch);
        else rc = queue_round(nv - 1, ch, nullptr, nullptr);
    }
    if (!rc && tail_queued) {
        // the gather and the copy wait behind the index gate (queued with the final fold)
        for (int q = 0; q < nq; q++) pin_idx[q] = mam::tr_challenge_scalar(t, nullptr, 0) % (u64)(cw_len / 2);
        mam_gate_open(idx_gate, nullptr, 0);
        if (!copy_queued) rc = queue_copy();
    } else if (!rc) {
        QueryIdx qin;
        if (nq > 0) {
            std::vector<u64> idx(nq);
            for (int q = 0; q < nq; q++) idx[q] = mam::tr_challenge_scalar(t, nullptr, 0) % (u64)(cw_len / 2);
            const bool inl = nq <= QIDX_INLINE;
            if (inl)
                for (int q = 0; q < nq; q++) qin.v[q] = idx[q];
            else if (hipMemcpyAsync(d_idx, idx.data(), nq * sizeof(u64), hipMemcpyHostToDevice, s) != hipSuccess)
                rc = -1;
            if (!rc) rc = queue_gather(inl ? nullptr : d_idx, qin);
        } else {
            rc = queue_gather(nullptr, qin);
        }
        if (!rc) rc = queue_copy();
    }
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = -1;
    // a failure may leave a queued round behind its gate: release it so the stream drains
    if (rc && pq) mam_gate_open_all();
    if (pq && mam_gate_take_failure() && !rc) rc = -1;  // a gate gave up: stale challenges
    return rc;
}

int open_ext2(const Commitment &C, const u64 *point2, const u64 *d_table, int table_nv, mam_transcript *t, int nq,
              u64 *proof) {
    const int nv = C.nv;
    if (nv < 2 || table_nv < nv) return -2;
    hipStream_t s = mam_stream();
    mam_stream_pool_setup();
    const size_t n = (size_t)1 << nv;
    // the transcript preamble (root, point) is absorbed by open_core once the first launches
    // are queued
    // rounds whose eq / bh fold runs on a launch of its own take the product form (no eq
    // table); from the first round that rides along a lane-parallel tree launch on, the
    // scaled eq table is built and folded as usual
    int K = nv - 1;
    if (C.hash == MAM_HASH_MONOLITH) {
        for (int r = 0; r < nv - 1; r++) {
            if (((size_t)1 << (nv + C.lr - r - 2)) < mam_merkle_narrow_min()) {
                K = r;
                break;
            }
        }
    }
    // Product-form openings run on the commitment's bit-reversed bh with the point's
    // coordinates reversed (OpenSrc::rev; MAM_OPEN_REV=0: reverse bh to natural order first).
    const bool rev = K > 0 && open_rev();
    // point (and its reversed coordinates), natural-order bh (unless rev) and (K == 0 only)
    // the eq table in the thread's arena slot 3
    auto words = [](size_t w) { return (w + 31) & ~(size_t)31; };
    const size_t eq_words = K == 0 ? words(2 * n) : 0;
    u64 *arena = mam_dev_arena(words(4 * nv) + (rev ? 0 : words(n)) + eq_words, 3);
    if (!arena) return -1;
    u64 *d_point = arena, *d_pt_rev = arena + 2 * nv, *bh = arena + words(4 * nv);
    u64 *eq = K == 0 ? bh + words(n) : nullptr;
    // the point goes through the thread's pinned staging words (a copy from pageable memory
    // would be staged and waited for by the runtime)
    u64 *stage = pinned_scratch();
    if (!stage || 4 * nv > 128) return -1;
    memcpy(stage + 64, point2, 2 * nv * sizeof(u64));
    for (int i = 0; i < nv; i++) {
        stage[64 + 2 * nv + 2 * i] = point2[2 * (nv - 1 - i)];
        stage[64 + 2 * nv + 2 * i + 1] = point2[2 * (nv - 1 - i) + 1];
    }
    if (hipMemcpyAsync(d_point, stage + 64, 4 * nv * sizeof(u64), hipMemcpyHostToDevice, s) != hipSuccess) return -1;
    int rc = K == 0 ? mam_eq_table(d_point, nv, eq, true, s) : 0;
    if (!rc && !rev) rc = mam_bitrev(C.bh_evals, bh, nv, 1, s);
    if (rc) return rc;
    OpenSrc S{K == 0 ? eq : nullptr, rev ? C.bh_evals : bh, true, C.codeword, true, C.tree, nv, C.lr, C.hash};
    S.pre_root = C.root;
    S.pre_point = point2;
    if (K > 0) {
        S.pt = d_point;
        S.pt_host = point2;
        S.K = K;
        S.rev = rev;
        S.pt_rev = d_pt_rev;
    }
    return open_core(S, d_table, table_nv, t, nq, proof);
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

// ---------------------------------------------------------------------------
// Batch opening (batch_open_ext2, basefold.rs:1202-1525; batch_verify_ext2 :1931-2221)
// ---------------------------------------------------------------------------
// Proof layout: outer_oracles[nv][3][2] | combined_root[4] | [open section of the combined
// codeword: inner_eval[2] | inner_oracles[nv][3][2] | folded_roots[nv-1][4] | final_cw |
// queries with Ext2 values at level 0] | nq x num_evals x { values[2] | path[nv+lr-1][4] }.
struct EvalClaim {
    int poly, point;
    u64 value[2];
};

long batch_proof_len(int nv, int lr, int nq, int ne) {
    return 6L * nv + 4 + proof_len(nv, lr, nq) + (long)nq * ne * (2 + 4L * (nv + lr - 1));
}

int batch_ell(int ne) {
    if (ne <= 1) return 1;
    int l = 0;
    while ((1 << l) < ne) l++;
    return l;
}

// eq_poly_host (basefold.rs:3008-3025): weights[i] = eq(t, bits of i), t_0 on the most
// significant index bit.
std::vector<u64> eq_weights_msb(const std::vector<u64> &tch) {
    std::vector<u64> w((size_t)1 << tch.size(), 0);
    w[0] = 1;
    for (size_t i = 0; i < tch.size(); i++) {
        const long half = 1L << i;
        const u64 omr = gl_sub(1, tch[i]);
        for (long j = half - 1; j >= 0; j--) {
            const u64 v = w[j];
            w[2 * j] = gl_mul(v, omr);
            w[2 * j + 1] = gl_mul(v, tch[i]);
        }
    }
    return w;
}

// eq_eval_host_ext2 (basefold.rs:3218-3230)
ext2 eq_eval_ext2(const std::vector<ext2> &r, const u64 *pt2) {
    ext2 res = e2_one();
    for (size_t i = 0; i < r.size(); i++) {
        const ext2 p = ld2(pt2 + 2 * i);
        res = e2_mul(res, e2_add(e2_mul(r[i], p), e2_mul(e2_sub(e2_one(), r[i]), e2_sub(e2_one(), p))));
    }
    return res;
}

const int LIN_MAX = 16;
struct LinSrc {
    const u64 *p[LIN_MAX];
    u64 s0[LIN_MAX], s1[LIN_MAX];
    int count;
};

// out[y] = (accumulate ? out[y] : 0) + sum_k s_k src_k[y] over base-field sources; the output
// is base (EXT_OUT false, s_k base) or Ext2 (s_k Ext2).
template <bool EXT_OUT>
__global__ void k_lincomb(LinSrc L, size_t len, int accumulate, u64 *__restrict__ out) {
    const size_t y = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (y >= len) return;
    if (EXT_OUT) {
        ext2 acc = accumulate ? e2_load(out, y) : e2_zero();
        for (int k = 0; k < L.count; k++) acc = e2_add(acc, e2_scale(gl_canon(L.p[k][y]), ext2{L.s0[k], L.s1[k]}));
        e2_store(out, y, acc);
    } else {
        u64 acc = accumulate ? out[y] : 0;
        for (int k = 0; k < L.count; k++) acc = gl_add(acc, gl_mul(gl_canon(L.p[k][y]), L.s0[k]));
        out[y] = acc;
    }
}

// Sum of `n` oracles (6 words each) into out (one wave).
__global__ void k_sum_oracles(const u64 *__restrict__ res, int n, u64 *__restrict__ out) {
    if (threadIdx.x >= 6) return;
    u64 acc = 0;
    for (int i = 0; i < n; i++) acc = gl_add(acc, res[6 * i + threadIdx.x]);
    out[threadIdx.x] = acc;
}

// Individual openings: block b = q * ne + e writes the base pair at the query's leaf and the
// authentication path in that commitment's tree. srcs[2e] = codeword, srcs[2e + 1] = tree.
__global__ void k_gather_individual(const u64 *const *__restrict__ srcs, const u64 *__restrict__ idx, int ne,
                                    int depth, u64 *__restrict__ out) {
    const int b = blockIdx.x;
    const int q = b / ne, e = b % ne;
    const u64 leaf = idx[q];
    const u64 *cw = srcs[2 * e], *tr = srcs[2 * e + 1];
    u64 *o = out + (size_t)b * (2 + 4 * (size_t)depth);
    if (threadIdx.x == 0) {
        o[0] = gl_canon(cw[2 * leaf]);
        o[1] = gl_canon(cw[2 * leaf + 1]);
    }
    const u64 leaves = (u64)1 << depth;
    for (int l = threadIdx.x; l < depth; l += blockDim.x) {
        const u64 lvl_off = 2 * leaves - 2 * (leaves >> l);
        const u64 *d = tr + 4 * (lvl_off + ((leaf >> l) ^ 1));
        u64 *w = o + 2 + 4 * l;
        w[0] = d[0];
        w[1] = d[1];
        w[2] = d[2];
        w[3] = d[3];
    }
}

// Launch k_lincomb over `srcs` in chunks of LIN_MAX (the first chunk overwrites `out`).
int lincomb(const std::vector<const u64 *> &srcs, const std::vector<ext2> &sc, bool ext_out, size_t len, u64 *out,
            hipStream_t s) {
    size_t k0 = 0;
    do {
        LinSrc L;
        memset(&L, 0, sizeof(L));
        L.count = (int)std::min((size_t)LIN_MAX, srcs.size() - k0);
        for (int k = 0; k < L.count; k++) {
            L.p[k] = srcs[k0 + k];
            L.s0[k] = sc[k0 + k].c0;
            L.s1[k] = sc[k0 + k].c1;
        }
        if (ext_out) k_lincomb<true><<<grid_1d(len), MAM_BLOCK, 0, s>>>(L, len, k0 > 0, out);
        else k_lincomb<false><<<grid_1d(len), MAM_BLOCK, 0, s>>>(L, len, k0 > 0, out);
        if (hipGetLastError() != hipSuccess) return -1;
        k0 += L.count;
    } while (k0 < srcs.size());
    return 0;
}

int batch_open_ext2(const std::vector<const Commitment *> &C, const u64 *points, int np, const EvalClaim *ev, int ne,
                    const u64 *d_table, int table_nv, mam_transcript *t, int nq, u64 *proof) {
    const int nv = C[0]->nv, lr = C[0]->lr, hash = C[0]->hash;
    if (nv < 2 || table_nv < nv || np < 1 || ne < 1) return -2;
    for (const Commitment *c : C)
        if (c->nv != nv || c->lr != lr || c->hash != hash) return -2;
    for (int e = 0; e < ne; e++)
        if (ev[e].poly < 0 || ev[e].poly >= (int)C.size() || ev[e].point < 0 || ev[e].point >= np) return -2;
    hipStream_t s = mam_stream();
    mam_stream_pool_setup();
    const size_t n = (size_t)1 << nv, cw_len = (size_t)1 << (nv + lr);
    // transcript preamble: roots, points, claimed values; batching weights
    for (const Commitment *c : C) observe_hash(t, c->root);
    for (int p = 0; p < np; p++)
        for (int i = 0; i < nv; i++) observe_ext2(t, points + 2L * (p * nv + i));
    for (int e = 0; e < ne; e++) observe_ext2(t, ev[e].value);
    std::vector<u64> tch(batch_ell(ne));
    for (auto &x : tch) x = mam::tr_challenge_scalar(t, nullptr, 0);
    const std::vector<u64> w = eq_weights_msb(tch);

    // device buffers in the thread's arena slot 2
    auto words = [](size_t x) { return (x + 31) & ~(size_t)31; };
    const size_t per_point = words(2 * n) + 2 * words(n) + 2 * (words(n) + words(n / 2));
    const size_t need = words(2L * np * nv) + np * per_point + words(RED_BLOCKS * 6 + 6) + words(8) +
                        words(6L * np) + words(6) + words(2 * cw_len) + 2 * words(2 * n) + words((cw_len - 1) * 4) +
                        words(2 * nv) + words(2 * n) + words(2L * ne) + words(nq) +
                        words((size_t)nq * ne * (2 + 4 * (nv + lr - 1)));
    u64 *arena = mam_dev_arena(need, 2);
    if (!arena) return -1;
    size_t used = 0;
    auto alloc = [&](size_t x) {
        u64 *p = arena + used;
        used += words(x);
        return p;
    };
    u64 *d_points = alloc(2L * np * nv);
    std::vector<u64 *> eqs(np), g1(np), gn(np), eqp0(np), eqp1(np), gp0(np), gp1(np);
    for (int p = 0; p < np; p++) {
        eqs[p] = alloc(2 * n);
        g1[p] = alloc(n);
        gn[p] = alloc(n);
        eqp0[p] = alloc(n);
        eqp1[p] = alloc(n / 2);
        gp0[p] = alloc(n);
        gp1[p] = alloc(n / 2);
    }
    u64 *red = alloc(RED_BLOCKS * 6 + 6), *cnt = alloc(8), *res = alloc(6L * np), *sum6 = alloc(6);
    u64 *ccw = alloc(2 * cw_len), *cbh1 = alloc(2 * n), *cbhn = alloc(2 * n), *ctree = alloc((cw_len - 1) * 4);
    u64 *d_r = alloc(2 * nv), *eqr = alloc(2 * n);
    const u64 **d_srcs = (const u64 **)alloc(2L * ne);
    u64 *d_idx = alloc(nq), *d_ind = alloc((size_t)nq * ne * (2 + 4 * (nv + lr - 1)));
    u64 *h = pinned_scratch();
    if (!h) return -1;
    int rc = 0;
    if (hipMemsetAsync(cnt, 0, 64, s) != hipSuccess) return -1;
    if (hipMemcpyAsync(d_points, points, 2L * np * nv * sizeof(u64), hipMemcpyHostToDevice, s) != hipSuccess)
        return -1;
    // g_p = sum over the claims at point p of w_e bh_e (natural order), eq_p
    for (int p = 0; p < np && !rc; p++) {
        std::vector<const u64 *> srcs;
        std::vector<ext2> sc;
        for (int e = 0; e < ne; e++)
            if (ev[e].point == p) {
                srcs.push_back(C[ev[e].poly]->bh_evals);
                sc.push_back(e2_base(w[e]));
            }
        rc = lincomb(srcs, sc, false, n, g1[p], s);
        if (!rc) rc = mam_bitrev(g1[p], gn[p], nv, 1, s);
        if (!rc) rc = mam_eq_table(d_points + 2L * p * nv, nv, eqs[p], true, s);
    }
    if (rc) return rc;
    // outer sumcheck of sum_p eq_p g_p: per-point oracles, summed on the device
    u64 *outer = proof;
    std::vector<ext2> r(nv);
    auto publish_sum = [&](u64 *dst) {
        k_sum_oracles<<<1, 64, 0, s>>>(res, np, sum6);
        if (hipGetLastError() != hipSuccess || read_back(h, sum6, 6, s)) return -1;
        memcpy(dst, h, 6 * sizeof(u64));
        for (int k = 0; k < 3; k++) observe_ext2(t, dst + 2 * k);
        return 0;
    };
    for (int p = 0; p < np && !rc; p++) rc = mam_open_oracle0(eqs[p], gn[p], true, n / 2, RedScratch{red, (u32 *)cnt, res + 6L * p}, s);
    if (!rc) rc = publish_sum(outer);
    std::vector<const u64 *> ceq(eqs.begin(), eqs.end()), cg(gn.begin(), gn.end());
    size_t sc_size = n;
    for (int round = 0; round < nv - 1 && !rc; round++) {
        r[round] = sample_ext2(t);
        for (int p = 0; p < np && !rc; p++) {
            u64 *eo = (round & 1) ? eqp1[p] : eqp0[p], *go = (round & 1) ? gp1[p] : gp0[p];
            rc = mam_open_fold_oracle(ceq[p], cg[p], round == 0, r[round], eo, go, sc_size / 4,
                                      RedScratch{red, (u32 *)cnt, res + 6L * p}, s);
            ceq[p] = eo;
            cg[p] = go;
        }
        sc_size /= 2;
        if (!rc) rc = publish_sum(outer + 6L * (round + 1));
    }
    if (rc) return rc;
    r[nv - 1] = sample_ext2(t);
    // combined codeword and bh: sum_e eq(r, point_e) w_e (codeword_e, bh_e)
    std::vector<ext2> eqpt(np), sc(ne);
    for (int p = 0; p < np; p++) eqpt[p] = eq_eval_ext2(r, points + 2L * p * nv);
    std::vector<const u64 *> cws(ne), bhs(ne);
    for (int e = 0; e < ne; e++) {
        sc[e] = e2_scale(w[e], eqpt[ev[e].point]);
        cws[e] = C[ev[e].poly]->codeword;
        bhs[e] = C[ev[e].poly]->bh_evals;
    }
    rc = lincomb(cws, sc, true, cw_len, ccw, s);
    if (!rc) rc = lincomb(bhs, sc, true, n, cbh1, s);
    if (!rc) rc = mam_bitrev(cbh1, cbhn, nv, 2, s);
    // a combination of replicated codewords is replicated
    bool all_rep = true;
    for (int e = 0; e < ne; e++) all_rep = all_rep && C[ev[e].poly]->replicated;
    if (!rc) rc = mam_merkle_build_copies(ccw, ctree, cw_len / 2, hash, 1, all_rep ? 1 << lr : 1, s);
    if (!rc) rc = read_back(h, ctree + mam_merkle_root_offset(cw_len / 2), 4, s);
    if (rc) return rc;
    u64 *croot = proof + 6L * nv;
    memcpy(croot, h, 4 * sizeof(u64));
    observe_hash(t, croot);
    // inner commit phase on the combined codeword at r
    std::vector<u64> r2(2 * nv);
    for (int i = 0; i < nv; i++) {
        r2[2 * i] = r[i].c0;
        r2[2 * i + 1] = r[i].c1;
    }
    if (hipMemcpyAsync(d_r, r2.data(), 2 * nv * sizeof(u64), hipMemcpyHostToDevice, s) != hipSuccess) return -1;
    rc = mam_eq_table(d_r, nv, eqr, true, s);
    if (rc) return rc;
    u64 *inner = croot + 4;
    OpenSrc S{eqr, cbhn, false, ccw, false, ctree, nv, lr, hash};
    rc = open_core(S, d_table, table_nv, t, nq, inner);
    if (rc || nq == 0) return rc;
    // individual openings at the same leaves
    const long qlen = query_len(nv, lr);
    const u64 *queries = inner + proof_len(nv, lr, 0);
    std::vector<u64> idx(nq);
    for (int q = 0; q < nq; q++) idx[q] = queries[(size_t)q * qlen];
    std::vector<const u64 *> srcs(2L * ne);
    for (int e = 0; e < ne; e++) {
        srcs[2 * e] = C[ev[e].poly]->codeword;
        srcs[2 * e + 1] = C[ev[e].poly]->tree;
    }
    const int depth = nv + lr - 1;
    if (hipMemcpyAsync(d_srcs, srcs.data(), srcs.size() * sizeof(u64 *), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_idx, idx.data(), nq * sizeof(u64), hipMemcpyHostToDevice, s) != hipSuccess)
        return -1;
    k_gather_individual<<<nq * ne, 64, 0, s>>>(d_srcs, d_idx, ne, depth, d_ind);
    if (hipGetLastError() != hipSuccess) return -1;
    u64 *o_ind = inner + proof_len(nv, lr, nq);
    if (hipMemcpyAsync(o_ind, d_ind, (size_t)nq * ne * (2 + 4 * depth) * sizeof(u64), hipMemcpyDeviceToHost, s) !=
        hipSuccess)
        return -1;
    return hipStreamSynchronize(s) == hipSuccess ? 0 : -1;
}

bool check_transition(const u64 *o, const u64 *on, ext2 ch) {
    const ext2 val = e2_add(ld2(o), e2_add(e2_mul(ld2(o + 2), ch), e2_mul(ld2(o + 4), e2_mul(ch, ch))));
    const ext2 n0 = ld2(on);
    return e2_eq(val, e2_add(n0, e2_add(n0, e2_add(ld2(on + 2), ld2(on + 4)))));
}

int batch_verify_ext2(const u64 *roots, int nc, int nv, int lr, int hash, const u64 *points, int np,
                      const EvalClaim *ev, int ne, const u64 *proof, int nq, const u64 *table, int table_nv,
                      mam_transcript *t) {
    if (nv < 2 || table_nv < nv || nc < 1 || np < 1 || ne < 1) return 0;
    for (int e = 0; e < ne; e++)
        if (ev[e].poly < 0 || ev[e].poly >= nc || ev[e].point < 0 || ev[e].point >= np) return 0;
    const u64 *outer = proof, *croot = proof + 6L * nv, *pc = croot + 4;
    const ext2 inner_eval = ld2(pc);
    const u64 *inner = pc + 2, *froots = inner + 6L * nv, *final_cw = froots + 4L * (nv - 1);
    const u64 *queries = final_cw + 2L * (1L << lr);
    for (int c = 0; c < nc; c++) observe_hash(t, roots + 4L * c);
    for (int p = 0; p < np; p++)
        for (int i = 0; i < nv; i++) observe_ext2(t, points + 2L * (p * nv + i));
    for (int e = 0; e < ne; e++) observe_ext2(t, ev[e].value);
    std::vector<u64> tch(batch_ell(ne));
    for (auto &x : tch) x = mam::tr_challenge_scalar(t, nullptr, 0);
    const std::vector<u64> w = eq_weights_msb(tch);
    ext2 claimed = e2_zero();
    for (int e = 0; e < ne; e++) claimed = e2_add(claimed, e2_scale(w[e], ld2(ev[e].value)));
    const ext2 o0 = ld2(outer);
    if (!e2_eq(e2_add(o0, e2_add(o0, e2_add(ld2(outer + 2), ld2(outer + 4)))), claimed)) return 0;
    for (int k = 0; k < 3; k++) observe_ext2(t, outer + 2 * k);
    std::vector<ext2> r(nv), ri(nv);
    for (int round = 0; round < nv - 1; round++) {
        r[round] = sample_ext2(t);
        if (!check_transition(outer + 6L * round, outer + 6L * (round + 1), r[round])) return 0;
        for (int k = 0; k < 3; k++) observe_ext2(t, outer + 6L * (round + 1) + 2 * k);
    }
    r[nv - 1] = sample_ext2(t);
    {
        const u64 *o = outer + 6L * (nv - 1);
        const ext2 ch = r[nv - 1];
        const ext2 fin = e2_add(ld2(o), e2_add(e2_mul(ld2(o + 2), ch), e2_mul(ld2(o + 4), e2_mul(ch, ch))));
        if (!e2_eq(fin, inner_eval)) return 0;
    }
    observe_hash(t, croot);
    const ext2 i0 = ld2(inner);
    if (!e2_eq(e2_add(i0, e2_add(i0, e2_add(ld2(inner + 2), ld2(inner + 4)))), inner_eval)) return 0;
    for (int k = 0; k < 3; k++) observe_ext2(t, inner + 2 * k);
    for (int round = 0; round < nv - 1; round++) {
        ri[round] = sample_ext2(t);
        if (!check_transition(inner + 6L * round, inner + 6L * (round + 1), ri[round])) return 0;
        for (int k = 0; k < 3; k++) observe_ext2(t, inner + 6L * (round + 1) + 2 * k);
        observe_hash(t, froots + 4L * round);
    }
    ri[nv - 1] = sample_ext2(t);
    std::vector<ext2> sc(ne);
    for (int e = 0; e < ne; e++) sc[e] = e2_scale(w[e], eq_eval_ext2(r, points + 2L * ev[e].point * nv));
    const u64 cw_half = (u64)1 << (nv + lr - 1);
    const long qlen = query_len(nv, lr);
    const int depth = nv + lr - 1;
    const u64 *indiv = queries + (size_t)nq * qlen;
    for (int q = 0; q < nq; q++) {
        const u64 leaf = mam::tr_challenge_scalar(t, nullptr, 0) % cw_half;
        const u64 *qp = queries + (size_t)q * qlen;
        if (qp[0] != leaf) return 0;
        const u64 *vals = qp + 1, *path = vals + 4L * nv;
        u64 lh[4];
        hash_ext2_leaf(hash, vals, lh);
        if (!verify_path(hash, lh, path, depth, leaf, croot)) return 0;
        path += 4L * depth;
        u64 fidx = leaf;
        for (int rr = 0; rr < nv; rr++) {
            const u64 *lvl = table + 2 * table_level_offset(table_nv, lr, rr);
            const ext2 v0 = ld2(vals + 4L * rr), v1 = ld2(vals + 4L * rr + 2);
            const ext2 cx = e2_sub(ri[rr], e2_base(gl_canon(lvl[2 * fidx])));
            const ext2 fr = e2_add(v0, e2_mul(e2_mul(cx, e2_sub(v1, v0)), e2_base(gl_canon(lvl[2 * fidx + 1]))));
            if (rr < nv - 1) {
                const u64 *nx = vals + 4L * (rr + 1);
                if (!e2_eq(fr, (fidx & 1) ? ld2(nx + 2) : ld2(nx))) return 0;
                hash_ext2_leaf(hash, nx, lh);
                const int plen = nv + lr - 2 - rr;
                if (!verify_path(hash, lh, path, plen, fidx / 2, froots + 4L * rr)) return 0;
                path += 4L * plen;
            } else if (!e2_eq(fr, ld2(final_cw + 2 * fidx))) {
                return 0;
            }
            fidx /= 2;
        }
        ext2 cv0 = e2_zero(), cv1 = e2_zero();
        for (int e = 0; e < ne; e++) {
            const u64 *ip = indiv + ((size_t)q * ne + e) * (2 + 4 * depth);
            hash_gl_leaf(hash, ip[0], ip[1], lh);
            if (!verify_path(hash, lh, ip + 2, depth, leaf, roots + 4L * ev[e].poly)) return 0;
            cv0 = e2_add(cv0, e2_scale(gl_canon(ip[0]), sc[e]));
            cv1 = e2_add(cv1, e2_scale(gl_canon(ip[1]), sc[e]));
        }
        if (!e2_eq(cv0, ld2(vals)) || !e2_eq(cv1, ld2(vals + 2))) return 0;
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

MAM_API long mam_basefold_batch_proof_len(int num_vars, int log_rate, int num_queries, int num_evals) {
    return batch_proof_len(num_vars, log_rate, num_queries, num_evals);
}

// batch_open_ext2: handles of commitments with equal num_vars, log_rate and hash; points
// (host) [num_points][nv][2]; claims (host) with poly / point indices into those arrays.
MAM_API int mam_basefold_batch_open_ext2(void *const *handles, int num_comms, const u64 *points, int num_points,
                                         const EvalClaim *evals, int num_evals, const u64 *d_table, int table_nv,
                                         mam_transcript *t, int num_queries, u64 *proof) {
    if (!handles || num_comms < 1 || !points || !evals || !d_table || !t || !proof || num_queries < 0) return -2;
    std::vector<const Commitment *> C(num_comms);
    for (int c = 0; c < num_comms; c++) {
        C[c] = (const Commitment *)handles[c];
        if (!C[c]) return -2;
    }
    return batch_open_ext2(C, points, num_points, evals, num_evals, d_table, table_nv, t, num_queries, proof);
}

// Returns 1 if the batch proof verifies, 0 otherwise (host only; table is the host copy).
MAM_API int mam_basefold_batch_verify_ext2(const u64 *roots, int num_comms, int num_vars, int log_rate, int hash,
                                           const u64 *points, int num_points, const EvalClaim *evals, int num_evals,
                                           const u64 *proof, int num_queries, const u64 *table, int table_nv,
                                           mam_transcript *t) {
    if (!roots || !points || !evals || !proof || !table || !t) return 0;
    return batch_verify_ext2(roots, num_comms, num_vars, log_rate, hash, points, num_points, evals, num_evals, proof,
                             num_queries, table, table_nv, t);
}

// Returns 1 if the proof verifies, 0 otherwise (host only; table is the host copy).
MAM_API int mam_basefold_verify_ext2(const u64 *root4, const u64 *point2, int num_vars, int log_rate, int hash,
                                     const u64 *proof, int num_queries, const u64 *table, int table_nv,
                                     mam_transcript *t) {
    return verify_ext2(root4, point2, num_vars, log_rate, hash, proof, num_queries, table, table_nv, t);
}

// Host Monolith permutation (for checking the device one).
MAM_API void mam_monolith_permute_host(u64 *state12) { mono::permute(state12); }
