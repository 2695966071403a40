// Internal interface of the Merkle tree builder (k_merkle.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#define MAM_HASH_MONOLITH 0
#define MAM_HASH_POSEIDON2 1

// Build the full tree over `num_leaves` leaves (a power of two) on stream `s`, without
// synchronizing. Returns 0, -1 (launch failure) or -2 (bad argument).
int mam_merkle_build(const uint64_t *d_cw, uint64_t *d_tree, size_t num_leaves, int hash, int ext2_leaves,
                     hipStream_t s);

// The same tree for leaves that are `copies` (a power of two) repetitions of the first
// num_leaves / copies leaves, as over a Basefold commitment's codeword (2^log_rate identical
// blocks): hashes only the first copy's subtree and stores every digest at all copies.
int mam_merkle_build_copies(const uint64_t *d_cw, uint64_t *d_tree, size_t num_leaves, int hash, int ext2_leaves,
                            int copies, hipStream_t s);

// u64 offset of the root digest in a tree with `num_leaves` leaves.
size_t mam_merkle_root_offset(size_t num_leaves);
