// Opt-in timing of kernel groups with HIP events recorded on the launching stream.
// Disabled by default (no events, no cost); bench.py enables it to measure the dominant
// kernels inside its timed region. Each scope adds its algorithmic work (permutations
// or bytes) to its category.
#pragma once

#include <hip/hip_runtime.h>

enum MamProfCat {
    PROF_MERKLE = 0,
    PROF_OPEN_SUMCHECK = 1,
    PROF_COMMIT_TRANSFORM = 2,
    PROF_SUMCHECK = 3,
    PROF_LEAF_LAUNCH = 4,  // one k_tree_levels<Monolith> launch hashing base-field leaves
    PROF_AJTAI_GEMM = 5,   // one k_aj_gemm launch (work = matrix + limb-plane bytes)
    PROF_AJTAI_RING = 6,   // one dense ring-Ajtai launch (work = ring columns x witnesses)
    PROF_NCAT = 7
};

bool mam_prof_enabled(int cat);
// Suspends the timing on the calling thread (for launches recorded into a graph).
void mam_prof_suspend(bool on);
void mam_prof_push(int cat, double work, hipEvent_t a, hipEvent_t b);

struct ProfScope {
    int cat;
    double work;
    hipStream_t s;
    hipEvent_t a = nullptr;
    ProfScope(int c, double w, hipStream_t st) : cat(c), work(w), s(st) {
        if (mam_prof_enabled(c) && hipEventCreate(&a) == hipSuccess) {
            if (hipEventRecord(a, s) != hipSuccess) {
                (void)hipEventDestroy(a);
                a = nullptr;
            }
        }
    }
    ~ProfScope() {
        if (!a) return;
        hipEvent_t b;
        if (hipEventCreate(&b) == hipSuccess && hipEventRecord(b, s) == hipSuccess) {
            mam_prof_push(cat, work, a, b);
        } else {
            (void)hipEventDestroy(a);
        }
    }
};
