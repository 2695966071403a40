// Monolith-w12 permutation over q = 2^64 - 2^32 - 31, the hash of the zk-torch-4 transcript
// (zk-torch-4/src/crypto/monolith.rs:33-213): concrete(0), then six rounds of bars (the byte
// S-box on words 0..3), bricks (x_i += x_{i-1}^2 for i = 11 down to 1, i.e. from the old
// values) and concrete(r) = circulant MDS (first row 7, 23, 8, 26, 13, 10, 9, 7, 6, 22, 21, 8,
// the same small matrix as the Goldilocks Monolith) plus the round-r constants (rounds 1..5;
// rounds 0 and 6 add zero). Every word leaves canonical.
#pragma once

#include "fieldq.h"
#include "monolith.h"
#include "monolith_q_rc.h"

namespace monoq {

static constexpr int WIDTH = 12;

struct RcTable {
    u64 v[5][WIDTH];
};
static constexpr RcTable kRc = {{MONOQ_RC_ROWS}};

// y_r = sum_c M[r][c] x_c + rc_r: twelve 64 x 5-bit products, exact in 128 bits, one reduction
MAM_HD void concrete(u64 x[WIDTH], const u64 *rc) {
    u64 y[WIDTH];
    for (int r = 0; r < WIDTH; r++) {
        u64 lo = rc ? rc[r] : 0, hi = 0;
        for (int c = 0; c < WIDTH; c++) {
            u64 pl, ph;
            mul_wide(fq::canon(x[c]), mono::row0((c - r + WIDTH) % WIDTH), pl, ph);
            lo += pl;
            hi += ph + (lo < pl);
        }
        y[r] = fq::reduce128(lo, hi);
    }
    for (int r = 0; r < WIDTH; r++) x[r] = y[r];
}

MAM_HD void permute(u64 x[WIDTH]) {
    concrete(x, nullptr);
    for (int round = 1; round <= 6; round++) {
        for (int i = 0; i < 4; i++) x[i] = mono::bar(x[i]);
        for (int i = WIDTH - 1; i >= 1; i--) x[i] = fq::add(fq::canon(x[i]), fq::sqr(x[i - 1]));
        concrete(x, round <= 5 ? kRc.v[round - 1] : nullptr);
    }
}

}  // namespace monoq
