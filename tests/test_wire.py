"""Proof wire format: bincode 1.x bytes of zk-torch-3's SumcheckProof (host code of the
library, no GPU): a hand-written known answer from the bincode 1.x format (u64 little-endian
integers and Vec lengths, structs as their fields in order, newtypes as their content), the
length formula, canonical residues, and agreement with the oracle's restatement."""
import numpy as np

import pyoracle as O

from morpheus_amd.prover import fingerprint, sumcheck_proof_bincode

P = O.P


def test_known_answer():
    proof = {"final_eval": np.array([1, 2], dtype=np.uint64),
             "round_msgs": np.array([[[3, 4], [5, P + 6]]], dtype=np.uint64)}
    want = bytes.fromhex(
        "0100000000000000" "0200000000000000"  # final_eval c0, c1
        "0100000000000000"                      # one round
        "0200000000000000"                      # two entries
        "0300000000000000" "0400000000000000"
        "0500000000000000" "0600000000000000")  # p + 6 written as its residue 6
    assert sumcheck_proof_bincode(proof) == want


def test_matches_oracle_restatement_and_fingerprint():
    rng = np.random.default_rng(8)
    msgs = rng.integers(0, 1 << 64, (20, 3, 2), dtype=np.uint64)
    fe = rng.integers(0, 1 << 64, 2, dtype=np.uint64)
    b = sumcheck_proof_bincode({"final_eval": fe, "round_msgs": msgs})
    assert len(b) == 16 + 8 + 20 * (8 + 3 * 16)
    assert b == O.sumcheck_proof_bincode(fe, msgs)
    assert len(fingerprint(b)) == 16 and fingerprint(b) == fingerprint(O.sumcheck_proof_bincode(fe, msgs))
    empty = sumcheck_proof_bincode({"final_eval": fe, "round_msgs": np.zeros((0, 3, 2), dtype=np.uint64)})
    assert len(empty) == 24 and empty[16:] == bytes(8)
