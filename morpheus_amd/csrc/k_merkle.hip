// Merkle trees over Basefold codewords with Monolith-w12 (the reference default) or
// Poseidon2-w8, and the Poseidon2 batch primitives.
//
// Tree layout (goldilocks-cuda-rs/src/merkle.rs:1-25): one buffer holding
// [level 0: N digests][level 1: N/2]...[root], 4 u64 per digest, (2N - 1) * 4 u64 in all.
// Leaf i of a base-field codeword hashes (cw[2i], cw[2i+1]); of an Ext2 codeword the four
// words cw[4i..4i+4) (wrapper.cu:1604-1648, 1909-1932). Node = compress(left, right).
//
// MI355X design: the reference launches one kernel and synchronizes per tree level
// (wrapper.cu:1653-1672, 1955-1988). Here one launch hashes a 256-wide slab of a level and
// then keeps compressing inside LDS while every wave of the workgroup is still full
// (256 -> 128 -> 64 nodes), so a 2^22-leaf tree is built by 8 launches with no host
// synchronization in between, and each digest is written to HBM exactly once.
//
// Replicated codewords. A Basefold commitment's codeword is 2^lr identical copies of a
// 2^nv-word block: the reference's encode repeats every coefficient 2^lr times, its
// butterflies act only on the bits above the low lr ones, and the bit reversal to Type1
// order moves those repetition bits to the top (goldilocks-cuda-rs/cuda/wrapper.cu:1277-1312,
// cuda/basefold.cuh:157-205). Leaf i and leaf i + L / 2^lr then hash the same words, so
// every level of the tree is 2^lr copies of the subtree over the first copy's leaves, and
// the top lr levels compress a node with itself. Builders given `rep` = 2^lr hash the first
// copy's subtree once, store each digest at all rep positions of its level, and form the
// top levels from the subtree root (k_tree_self_top): the same tree with half the
// permutations at rate 1/2.
#include <stdlib.h>

#include "basefold.h"
#include "common.h"
#include "merkle.h"
#include "monolith.h"
#include "monolith_wide.h"
#include "open_fold.h"
#include "poseidon2.h"
#include "profile.h"
#include "sync.h"

struct MonoHash {
    static constexpr bool kWide = true;
    static MAM_HD void gl_leaf(u64 a, u64 b, u64 *o) { mono::hash_gl_leaf(a, b, o); }
    static MAM_HD void ext2_leaf(const u64 *ab, u64 *o) { mono::hash_ext2_leaf(ab, o); }
    static MAM_HD void node(const u64 *l, const u64 *r, u64 *o) { mono::compress(l, r, o); }
};
struct P2Hash {
    static constexpr bool kWide = false;
    static MAM_HD void gl_leaf(u64 a, u64 b, u64 *o) { p2::hash_gl_leaf(a, b, o); }
    static MAM_HD void ext2_leaf(const u64 *ab, u64 *o) { p2::hash_ext2_leaf(ab, o); }
    static MAM_HD void node(const u64 *l, const u64 *r, u64 *o) { p2::compress(l, r, o); }
};

static constexpr int TREE_BLOCK = 256;
static constexpr unsigned UJ_BLOCKS = 256;  // workgroups of a product-form job riding along a tree launch

// Source kinds for the first level a launch writes: codeword leaves (base or Ext2),
// children nodes, or leaves of a codeword folded on the fly (base or Ext2 input).
enum { SRC_GL = 0, SRC_EXT2 = 1, SRC_NODES = 2, SRC_FOLD_GL = 3, SRC_FOLD_EXT2 = 4 };

struct FoldSrc {
    const u64 *table;  // (x0, w) pairs of this fold level
    ext2 alpha;
    u64 *out;          // folded codeword (Ext2)
    const u64 *alpha_p;  // non-null: the challenge is read here (written on the device after launch)
};

__device__ __forceinline__ ext2 fold_alpha(const FoldSrc &fs) {
    return fs.alpha_p ? ext2{fs.alpha_p[0], fs.alpha_p[1]} : fs.alpha;
}

// Writes `nlev` consecutive levels. The first written level has `n_first` computed nodes
// (rep copies of them make the level) and starts at digest index `first_off` of `tree`;
// each next level follows the previous one. Node k of a level is stored at k + c * n for
// every copy c < rep, n being the level's computed node count.
// ujb > 0 (folded-codeword launches only): the first ujb workgroups run the round's
// product-form bh fold and U sums (uj, open_fold.h) instead of tree work; the tree part is
// the rest of the grid.
template <class H, int SRC>
__global__ __launch_bounds__(TREE_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_tree_levels(const u64 *__restrict__ src, u64 *__restrict__ tree,
                                                            size_t first_off, size_t n_first, int nlev, FoldSrc fs,
                                                            int rep, UJob uj, unsigned ujb) {
    __shared__ u64 sh[TREE_BLOCK * 4];
    const int t = threadIdx.x;
    if constexpr (SRC == SRC_FOLD_GL || SRC == SRC_FOLD_EXT2) {
        if (blockIdx.x < ujb) {  // workgroup-uniform
            open_u_run<SRC == SRC_FOLD_GL, true, true>(uj, blockIdx.x, ujb);
            return;
        }
    }
    const unsigned tb = blockIdx.x - ujb;
    const size_t i = (size_t)tb * TREE_BLOCK + t;
    u64 d[4];
    if (i < n_first) {
        if (SRC == SRC_GL) {
            H::gl_leaf(src[2 * i], src[2 * i + 1], d);
        } else if (SRC == SRC_EXT2) {
            u64 ab[4] = {src[4 * i], src[4 * i + 1], src[4 * i + 2], src[4 * i + 3]};
            H::ext2_leaf(ab, d);
        } else if (SRC == SRC_NODES) {
            H::node(src + 8 * i, src + 8 * i + 4, d);
        } else {
            // folded elements 2i and 2i+1 come from input pairs (4i, 4i+1) and (4i+2, 4i+3)
            ext2 f[2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const size_t k = 2 * i + h;
                ext2 v0, v1;
                if (SRC == SRC_FOLD_GL) {
                    v0 = e2_base(gl_canon(src[2 * k]));
                    v1 = e2_base(gl_canon(src[2 * k + 1]));
                } else {
                    v0 = e2_load(src, 2 * k);
                    v1 = e2_load(src, 2 * k + 1);
                }
                const u64 x0 = gl_canon(fs.table[2 * k]), w = gl_canon(fs.table[2 * k + 1]);
                f[h] = mam_fold_value(v0, v1, e2_sub(fold_alpha(fs), e2_base(x0)), w);
                e2_store(fs.out, k, f[h]);
            }
            u64 ab[4] = {f[0].c0, f[0].c1, f[1].c0, f[1].c1};
            H::ext2_leaf(ab, d);
        }
        for (int c = 0; c < rep; c++) {
            u64 *o = tree + 4 * (first_off + (size_t)c * n_first + i);
#pragma unroll
            for (int k = 0; k < 4; k++) o[k] = d[k];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) sh[4 * t + k] = d[k];
    }
    size_t level_off = first_off;
    size_t level_n = n_first;
    int block_n = (int)(n_first < TREE_BLOCK ? n_first : TREE_BLOCK);
    for (int l = 1; l < nlev; l++) {
        mam_barrier();
        level_off += level_n * rep;
        level_n >>= 1;
        block_n >>= 1;
        u64 r[4];
        if (t < block_n) H::node(sh + 8 * t, sh + 8 * t + 4, r);
        mam_barrier();
        if (t < block_n) {
            for (int c = 0; c < rep; c++) {
                u64 *o = tree + 4 * (level_off + (size_t)c * level_n + (size_t)tb * (TREE_BLOCK >> l) + t);
#pragma unroll
                for (int k = 0; k < 4; k++) o[k] = r[k];
            }
#pragma unroll
            for (int k = 0; k < 4; k++) sh[4 * t + k] = r[k];
        }
    }
}

// ---------------------------------------------------------------------------
// Lane-parallel variant for small levels: 16 lanes per permutation (monolith_wide.h),
// WIDE_GROUPS permutations per pass; a workgroup writes a level slab and then the
// in-LDS levels above it down to one node per workgroup.
// ---------------------------------------------------------------------------
static constexpr int WIDE_GROUPS = 32;

// The last oj.blocks workgroups of a launch run the round's eq / bh fold and oracle job
// (open_fold.h) instead of tree work; the tree part is the first gridDim.x - oj.blocks.
// mj.box (only in a one-workgroup launch that ends at the root): post mj.a's words and the
// root to the host mailbox once the root is written.
template <int SRC>
__global__ __launch_bounds__(WIDE_GROUPS * 16) void k_tree_wide(const u64 *__restrict__ src, u64 *__restrict__ tree,
                                                                size_t first_off, size_t n_first, int nlev,
                                                                FoldSrc fs, OracleJob oj, MailJob mj, int rep) {
    __shared__ u64 sh[WIDE_GROUPS * 4];
    __shared__ __align__(16) u64 wst[WIDE_GROUPS][16];  // per-group permutation slots
    if constexpr (SRC == SRC_FOLD_GL || SRC == SRC_FOLD_EXT2) {  // only folded trees carry the job
        if (oj.blocks) {
            const unsigned tree_blocks = gridDim.x - oj.blocks;
            if (blockIdx.x >= tree_blocks) {  // workgroup-uniform
                if (oj.bh_base)
                    open_fold_oracle_run<true, true>(oj, blockIdx.x - tree_blocks, oj.blocks, sh);
                else
                    open_fold_oracle_run<false, true>(oj, blockIdx.x - tree_blocks, oj.blocks, sh);
                return;
            }
        }
    }
    const int g = threadIdx.x >> 4, j = threadIdx.x & 15;
    mono_wide::LaneConsts lc;
    mono_wide::lane_consts(j, lc);
    const size_t i = (size_t)blockIdx.x * WIDE_GROUPS + g;
    const bool active = i < n_first;
    u64 x = 0;
    if (active) {
        if (SRC == SRC_GL) {
            if (j < 2) x = src[2 * i + j];
        } else if (SRC == SRC_EXT2) {
            if (j < 4) x = src[4 * i + j];
        } else if (SRC == SRC_NODES) {
            if (j < 8) x = src[8 * i + j];
        } else if (j < 4) {
            // lanes 0,1 build folded element 2i, lanes 2,3 element 2i+1; lane parity picks c0/c1
            const size_t k = 2 * i + (j >> 1);
            ext2 v0, v1;
            if (SRC == SRC_FOLD_GL) {
                v0 = e2_base(gl_canon(src[2 * k]));
                v1 = e2_base(gl_canon(src[2 * k + 1]));
            } else {
                v0 = e2_load(src, 2 * k);
                v1 = e2_load(src, 2 * k + 1);
            }
            const u64 x0 = gl_canon(fs.table[2 * k]), w = gl_canon(fs.table[2 * k + 1]);
            const ext2 f = mam_fold_value(v0, v1, e2_sub(fold_alpha(fs), e2_base(x0)), w);
            x = (j & 1) ? f.c1 : f.c0;
            fs.out[2 * k + (j & 1)] = x;
        }
    }
    x = mono_wide::permute(x, j, wst[g], lc);
    if (active && j < 4) {
        for (int c = 0; c < rep; c++) tree[4 * (first_off + (size_t)c * n_first + i) + j] = x;
        sh[4 * g + j] = x;
    }
    size_t level_off = first_off, level_n = n_first;
    int block_n = (int)(n_first < WIDE_GROUPS ? n_first : WIDE_GROUPS);
    for (int l = 1; l < nlev; l++) {
        mam_barrier();
        level_off += level_n * rep;
        level_n >>= 1;
        block_n >>= 1;
        u64 y = (g < block_n && j < 8) ? sh[8 * g + j] : 0;
        mam_barrier();
        // wave-uniform: a wave (4 groups) runs the permutation if its first group is active
        if ((g & ~3) < block_n) y = mono_wide::permute(y, j, wst[g], lc);
        if (g < block_n && j < 4) {
            for (int c = 0; c < rep; c++)
                tree[4 * (level_off + (size_t)c * level_n + (size_t)blockIdx.x * (WIDE_GROUPS >> l) + g) + j] = y;
            sh[4 * g + j] = y;
        }
    }
    if (mj.box) {
        mam_barrier();  // the root is in sh[0..3]
        const int t = threadIdx.x;
        if (t < mj.na + 4) mj.box[t] = t < mj.na ? mj.a[t] : sh[t - mj.na];
        mam_mailbox_raise(mj.box, mj.seq);
        mam_gate_wait(mj.gate);  // the next round's challenge, when it is pre-queued
    }
}

// Top levels of a tree over rep copies of one subtree: the level at digest `prev_off` holds
// rep copies of the subtree root; each level above holds half as many copies of the
// compression of a node with itself, up to the root.
template <class H>
__global__ void k_tree_self_top(u64 *__restrict__ tree, size_t prev_off, int rep) {
    if (H::kWide) {
        // one 16-lane group
        __shared__ __align__(16) u64 wst[16];
        const int j = threadIdx.x & 15;
        mono_wide::LaneConsts lc;
        mono_wide::lane_consts(j, lc);
        u64 h = j < 8 ? tree[4 * prev_off + (j & 3)] : 0;
        size_t off = prev_off + rep;
        for (int n = rep >> 1; n >= 1; n >>= 1) {
            h = mono_wide::permute(h, j, wst, lc);
            if (j < 4)
                for (int c = 0; c < n; c++) tree[4 * (off + c) + j] = h;
            // the next level compresses this digest with itself: lanes 4..7 take lanes 0..3
            wst[j] = h;
            __builtin_amdgcn_wave_barrier();
            h = j < 8 ? wst[j & 3] : 0;
            __builtin_amdgcn_wave_barrier();
            off += n;
        }
    } else if (threadIdx.x == 0) {
        u64 h[4] = {tree[4 * prev_off], tree[4 * prev_off + 1], tree[4 * prev_off + 2], tree[4 * prev_off + 3]};
        size_t off = prev_off + rep;
        for (int n = rep >> 1; n >= 1; n >>= 1) {
            u64 r[4];
            H::node(h, h, r);
            for (int c = 0; c < n; c++)
                for (int k = 0; k < 4; k++) tree[4 * (off + c) + k] = r[k];
            for (int k = 0; k < 4; k++) h[k] = r[k];
            off += n;
        }
    }
}

static size_t narrow_min() {
    static const size_t v = [] {
        const char *e = getenv("MAM_MERKLE_NARROW_MIN");
        const size_t x = e ? (size_t)strtoull(e, nullptr, 10) : ((size_t)1 << 16);
        return x ? x : (size_t)1;
    }();
    return v;
}

size_t mam_merkle_narrow_min() { return narrow_min(); }

// Lane-parallel launches over more than one workgroup write at most this many levels
// (MAM_WIDE_LPL, default 6: down to one node per workgroup).
static int wide_lpl() {
    static const int v = [] {
        const char *e = getenv("MAM_WIDE_LPL");
        const int x = e ? atoi(e) : 6;
        return x >= 1 && x <= 6 ? x : 6;
    }();
    return v;
}

static int ilog2_sz(size_t n) {
    int l = 0;
    while ((size_t(1) << l) < n) l++;
    return l;
}

// Levels written by one launch whose first level has n nodes: all remaining levels when
// the whole rest fits one workgroup, else MAM_TREE_LPL levels (default 1), 1 below
// MAM_TREE_LPL_MIN nodes. A launch that goes on compressing in LDS keeps only half, then a
// quarter of its waves busy, and on the box two workgroups of a CU often put those waves on
// the same SIMDs; one level per launch measured faster at every size (2^22 Basefold proof:
// 3.21 ms with three levels per launch, 3.17 with two, 3.14 with one).
static int levels_per_launch(size_t n) {
    static const int lpl = [] {
        const char *e = getenv("MAM_TREE_LPL");
        const int v = e ? atoi(e) : 1;
        return v >= 1 && v <= 3 ? v : 1;
    }();
    static const size_t lpl_min = [] {
        const char *e = getenv("MAM_TREE_LPL_MIN");
        return e ? (size_t)strtoull(e, nullptr, 10) : (size_t)0;
    }();
    if (n <= (size_t)TREE_BLOCK) return ilog2_sz(n) + 1;
    return n < lpl_min ? 1 : lpl;
}

// oj (may be null): an opening round's fold + oracle job. It rides along the first launch
// when that launch is lane-parallel (workgroups of the same size), else it is launched on
// its own before the tree. mj (may be null): a mailbox post that the final launch makes
// after the root when that launch is one lane-parallel workgroup without the oracle job;
// *mail_posted tells the caller whether it did. rep (a power of two): the leaves are rep
// copies of the first num_leaves / rep (a replicated codeword); only that subtree is hashed.
// uj (may be null, folded-codeword sources only): an opening round's product-form job, run on
// extra workgroups of the first launch, which must be a one-lane launch.
template <class H>
static int build_tree(const u64 *d_cw, u64 *d_tree, size_t num_leaves, int leaf_src, const FoldSrc &fs,
                      hipStream_t s, const OracleJob *oj = nullptr, const MailJob *mj = nullptr,
                      bool *mail_posted = nullptr, int rep = 1, const UJob *uj = nullptr) {
    if (num_leaves == 0 || (num_leaves & (num_leaves - 1))) return -2;
    if (rep < 1 || (rep & (rep - 1)) || (size_t)rep > num_leaves) return -2;
    if (rep > 1 && (oj || mj || uj)) return -2;  // replicated trees are commitments: no opening jobs
    if (uj && (oj || (leaf_src != SRC_FOLD_GL && leaf_src != SRC_FOLD_EXT2) ||
               (H::kWide && num_leaves < narrow_min())))
        return -2;
    const UJob no_uj = {};
    const unsigned ujb = uj ? grid_stride(uj->pairs, TREE_BLOCK, UJ_BLOCKS) : 0;
    if (mail_posted) *mail_posted = false;
    const OracleJob none = {};
    const MailJob no_mail = {};
    OracleJob ride = none;
    if (oj) {
        if (H::kWide && num_leaves < narrow_min()) {
            ride = *oj;
            ride.blocks = grid_stride(oj->half_out, WIDE_GROUPS * 16, MAM_RED_GRID);
        } else if (mam_open_fold_oracle_job(*oj, s)) {
            return -1;
        }
    }
    // permutations executed: the subtree over one copy, then one per level above it
    ProfScope prof(PROF_MERKLE, 2.0 * (double)(num_leaves / rep) - 1.0 + ilog2_sz((size_t)rep), s);
    size_t n = num_leaves / rep, off = 0;  // n: computed nodes of the level being written
    int total_levels = ilog2_sz(n) + 1;
    int done = 0;
    bool first = true;
    const bool can_wide = H::kWide;
    while (done < total_levels) {
        if (can_wide && n < narrow_min()) {
            int nl = n <= (size_t)WIDE_GROUPS ? ilog2_sz(n) + 1 : ilog2_sz(WIDE_GROUPS) + 1;
            if (n > (size_t)WIDE_GROUPS && nl > wide_lpl()) nl = wide_lpl();
            if (nl > total_levels - done) nl = total_levels - done;
            unsigned grid = (unsigned)((n + WIDE_GROUPS - 1) / WIDE_GROUPS);
            const int threads = WIDE_GROUPS * 16;
            const OracleJob &j = first ? ride : none;
            // the final launch posts the mailbox when it is one workgroup and carries no job
            const bool post = mj && mj->box && done + nl == total_levels && grid == 1 && j.blocks == 0;
            const MailJob &m = post ? *mj : no_mail;
            if (first) {
                const unsigned g2 = grid + j.blocks;
                switch (leaf_src) {
                    case SRC_GL:
                        k_tree_wide<SRC_GL><<<g2, threads, 0, s>>>(d_cw, d_tree, 0, n, nl, fs, j, m, rep);
                        break;
                    case SRC_EXT2:
                        k_tree_wide<SRC_EXT2><<<g2, threads, 0, s>>>(d_cw, d_tree, 0, n, nl, fs, j, m, rep);
                        break;
                    case SRC_FOLD_GL:
                        k_tree_wide<SRC_FOLD_GL><<<g2, threads, 0, s>>>(d_cw, d_tree, 0, n, nl, fs, j, m, rep);
                        break;
                    default:
                        k_tree_wide<SRC_FOLD_EXT2><<<g2, threads, 0, s>>>(d_cw, d_tree, 0, n, nl, fs, j, m, rep);
                        break;
                }
            } else {
                // children: the computed part of the level that ends right before `off`
                const u64 *child = d_tree + 4 * (off - 2 * n * rep);
                k_tree_wide<SRC_NODES><<<grid, threads, 0, s>>>(child, d_tree, off, n, nl, fs, none, m, rep);
            }
            if (post && mail_posted) *mail_posted = true;
            if (hipGetLastError() != hipSuccess) return -1;
            for (int l = 0; l < nl; l++) {
                off += n * rep;
                n >>= 1;
            }
            done += nl;
            first = false;
            continue;
        }
        int nl = levels_per_launch(n);
        if (nl > total_levels - done) nl = total_levels - done;
        unsigned grid = (unsigned)((n + TREE_BLOCK - 1) / TREE_BLOCK);
        if (first) {
            // the single dominant kernel: base-field leaves plus in-LDS levels
            double launch_perms = 0;
            for (int l = 0; l < nl; l++) launch_perms += (double)(n >> l);
            ProfScope leaf_prof(leaf_src == SRC_GL && H::kWide ? PROF_LEAF_LAUNCH : -1, launch_perms, s);
            const UJob &u = uj ? *uj : no_uj;
            switch (leaf_src) {
                case SRC_GL:
                    k_tree_levels<H, SRC_GL><<<grid, TREE_BLOCK, 0, s>>>(d_cw, d_tree, 0, n, nl, fs, rep, no_uj, 0);
                    break;
                case SRC_EXT2:
                    k_tree_levels<H, SRC_EXT2><<<grid, TREE_BLOCK, 0, s>>>(d_cw, d_tree, 0, n, nl, fs, rep, no_uj, 0);
                    break;
                case SRC_FOLD_GL:
                    k_tree_levels<H, SRC_FOLD_GL><<<grid + ujb, TREE_BLOCK, 0, s>>>(d_cw, d_tree, 0, n, nl, fs, rep, u, ujb);
                    break;
                default:
                    k_tree_levels<H, SRC_FOLD_EXT2><<<grid + ujb, TREE_BLOCK, 0, s>>>(d_cw, d_tree, 0, n, nl, fs, rep, u, ujb);
                    break;
            }
        } else {
            // children: the computed part of the level that ends right before `off`
            const u64 *child = d_tree + 4 * (off - 2 * n * rep);
            k_tree_levels<H, SRC_NODES><<<grid, TREE_BLOCK, 0, s>>>(child, d_tree, off, n, nl, fs, rep, no_uj, 0);
        }
        if (hipGetLastError() != hipSuccess) return -1;
        for (int l = 0; l < nl; l++) {
            off += n * rep;
            n >>= 1;
        }
        done += nl;
        first = false;
    }
    if (rep > 1) {
        // the subtree roots are the rep nodes ending at `off`
        k_tree_self_top<H><<<1, 16, 0, s>>>(d_tree, off - rep, rep);
        if (hipGetLastError() != hipSuccess) return -1;
    }
    return 0;
}

int mam_merkle_build_copies(const u64 *d_cw, u64 *d_tree, size_t num_leaves, int hash, int ext2_leaves, int copies,
                            hipStream_t s) {
    FoldSrc fs = {nullptr, e2_zero(), nullptr};
    int src = ext2_leaves ? SRC_EXT2 : SRC_GL;
    if (hash == MAM_HASH_MONOLITH)
        return build_tree<MonoHash>(d_cw, d_tree, num_leaves, src, fs, s, nullptr, nullptr, nullptr, copies);
    if (hash == MAM_HASH_POSEIDON2)
        return build_tree<P2Hash>(d_cw, d_tree, num_leaves, src, fs, s, nullptr, nullptr, nullptr, copies);
    return -2;
}

int mam_merkle_build(const u64 *d_cw, u64 *d_tree, size_t num_leaves, int hash, int ext2_leaves, hipStream_t s) {
    return mam_merkle_build_copies(d_cw, d_tree, num_leaves, hash, ext2_leaves, 1, s);
}

int mam_fold_and_tree(const u64 *cw, bool base_cw, const u64 *table, ext2 alpha, u64 *fold_out, u64 *tree,
                      size_t folded_len, int hash, hipStream_t s, const OracleJob *oj, const MailJob *mj,
                      bool *mail_posted, const UJob *uj, const u64 *alpha_p) {
    if (mail_posted) *mail_posted = false;
    if (folded_len < 2) return -2;
    FoldSrc fs = {table, alpha, fold_out, alpha_p};
    int src = base_cw ? SRC_FOLD_GL : SRC_FOLD_EXT2;
    if (hash == MAM_HASH_MONOLITH)
        return build_tree<MonoHash>(cw, tree, folded_len / 2, src, fs, s, oj, mj, mail_posted, 1, uj);
    if (hash == MAM_HASH_POSEIDON2)
        return build_tree<P2Hash>(cw, tree, folded_len / 2, src, fs, s, oj, mj, mail_posted, 1, uj);
    return -2;
}

size_t mam_merkle_root_offset(size_t num_leaves) { return (2 * num_leaves - 2) * 4; }

// ---------------------------------------------------------------------------
// Reference-named tree entry points (ffi.rs:543-570): build and synchronize.
// ---------------------------------------------------------------------------
MAM_API int monolith_merkle_tree_gl_ffi(const u64 *d_codeword, u64 *d_tree, int num_leaves) {
    if (num_leaves <= 0) return -1;
    int rc = mam_merkle_build(d_codeword, d_tree, (size_t)num_leaves, MAM_HASH_MONOLITH, 0, mam_stream());
    return rc ? rc : mam_sync_status();
}
MAM_API int monolith_merkle_tree_ext2_ffi(const u64 *d_codeword, u64 *d_tree, int num_leaves) {
    if (num_leaves <= 0) return -1;
    int rc = mam_merkle_build(d_codeword, d_tree, (size_t)num_leaves, MAM_HASH_MONOLITH, 1, mam_stream());
    return rc ? rc : mam_sync_status();
}
MAM_API int poseidon2_merkle_tree_gl_ffi(const u64 *d_codeword, u64 *d_tree, int num_leaves) {
    if (num_leaves <= 0) return -1;
    int rc = mam_merkle_build(d_codeword, d_tree, (size_t)num_leaves, MAM_HASH_POSEIDON2, 0, mam_stream());
    return rc ? rc : mam_sync_status();
}
MAM_API int poseidon2_merkle_tree_ext2_ffi(const u64 *d_codeword, u64 *d_tree, int num_leaves) {
    if (num_leaves <= 0) return -1;
    int rc = mam_merkle_build(d_codeword, d_tree, (size_t)num_leaves, MAM_HASH_POSEIDON2, 1, mam_stream());
    return rc ? rc : mam_sync_status();
}

// ---------------------------------------------------------------------------
// Single permutations and Poseidon2 batches (wrapper.cu:864-928, 1990-2011)
// ---------------------------------------------------------------------------
__global__ void k_mono_permute_n(u64 *s, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u64 x[12];
#pragma unroll
    for (int k = 0; k < 12; k++) x[k] = s[12 * (size_t)i + k];
    mono::permute(x);
#pragma unroll
    for (int k = 0; k < 12; k++) s[12 * (size_t)i + k] = x[k];
}

MAM_API int monolith_permute_test_ffi(const u64 *h_input, u64 *h_output) {
    u64 *d = nullptr;
    if (hipMalloc(&d, 12 * sizeof(u64)) != hipSuccess) return -1;
    int rc = -1;
    if (hipMemcpyAsync(d, h_input, 96, hipMemcpyHostToDevice, mam_stream()) == hipSuccess) {
        k_mono_permute_n<<<1, 64, 0, mam_stream()>>>(d, 1);
        if (hipGetLastError() == hipSuccess &&
            hipMemcpyAsync(h_output, d, 96, hipMemcpyDeviceToHost, mam_stream()) == hipSuccess &&
            hipStreamSynchronize(mam_stream()) == hipSuccess)
            rc = 0;
    }
    hipFree(d);
    return rc;
}

// Monolith permutation of n states of 12 words in place (device pointers).
MAM_API int mam_monolith_permute_batch(u64 *d_states, int n) {
    if (n <= 0) return 0;
    k_mono_permute_n<<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(d_states, n);
    return mam_launch_status();
}

__global__ void k_p2_permute(const u64 *__restrict__ in, u64 *__restrict__ out, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u64 x[8];
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = in[8 * (size_t)i + k];
    p2::permute(x);
#pragma unroll
    for (int k = 0; k < 8; k++) out[8 * (size_t)i + k] = x[k];
}

__global__ void k_p2_compress(const u64 *__restrict__ l, const u64 *__restrict__ r, u64 *__restrict__ out, int n,
                              int interleaved) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u64 d[4];
    if (interleaved)
        p2::compress(l + 8 * (size_t)i, l + 8 * (size_t)i + 4, d);
    else
        p2::compress(l + 4 * (size_t)i, r + 4 * (size_t)i, d);
#pragma unroll
    for (int k = 0; k < 4; k++) out[4 * (size_t)i + k] = d[k];
}

MAM_API int poseidon2_hash_batch_ffi(const u64 *d_input, u64 *d_output, int n) {
    if (n <= 0) return 0;
    k_p2_permute<<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(d_input, d_output, n);
    return mam_launch_status();
}
MAM_API int poseidon2_compress_batch_ffi(const u64 *d_left, const u64 *d_right, u64 *d_output, int n) {
    if (n <= 0) return 0;
    k_p2_compress<<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(d_left, d_right, d_output, n, 0);
    return mam_launch_status();
}
MAM_API int poseidon2_merkle_layer_ffi(const u64 *d_input, u64 *d_output, int n) {
    if (n <= 0) return 0;
    k_p2_compress<<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(d_input, nullptr, d_output, n, 1);
    return mam_launch_status();
}

// Build a full tree from host-visible arguments on the calling thread's stream without
// synchronizing (for callers composing their own pipelines).
MAM_API int mam_merkle_tree_async(const u64 *d_codeword, u64 *d_tree, size_t num_leaves, int hash, int ext2_leaves) {
    return mam_merkle_build(d_codeword, d_tree, num_leaves, hash, ext2_leaves, mam_stream());
}
