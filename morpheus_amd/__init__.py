"""MI355X-native prover hot path: Goldilocks/Ext2 sumcheck, eq tables, Basefold with
Monolith/Poseidon2 Merkle trees, behind the goldilocks-cuda-rs C-ABI; the almost-Goldilocks
(q) crate's kernels, ring Ajtai and the zk-torch-4 transcript and sumcheck.

Importing requires the built HIP library (morpheus_amd/libmorpheus_amd.so); there is no
CPU fallback.
"""
from ._lib import EXPORTED, LIB_PATH, lib  # noqa: F401
from .device import DeviceBuffer, GpuError, PinnedArray, check, device_count, synchronize  # noqa: F401
from .prover import (HASH_MONOLITH, HASH_POSEIDON2, BasefoldCommitment, BasefoldTable,  # noqa: F401
                     GpuSumcheckStateExt2, Transcript, proof_len, split_proof, sumcheck_prove, sumcheck_verify,
                     verify_ext2)
from .qprover import QTranscript, monolith_q_permute, q_sumcheck_prove, q_sumcheck_verify  # noqa: F401

P = 0xFFFFFFFF00000001
