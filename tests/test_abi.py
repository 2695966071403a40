"""CPU tests of the C-ABI library: it loads without a GPU, exports every symbol the public
headers declare, and its host-side code (transcript, permutations, folding table,
verifier) agrees with the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import pyoracle as O
from refmath import GOLDEN, P, rand_field

import morpheus_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = set()
    for h in ("morpheus_gl.h", "morpheus_prover.h", "morpheus_agl.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|void|long|uint64_t)\s+\**\s*([a-z_0-9]+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_symbols_exported():
    names = header_functions()
    assert len(names) > 140
    out = subprocess.run(["nm", "-D", "--defined-only", M.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = names - exported
    assert not missing, missing
    for n in names:
        assert hasattr(M.lib, n)
    assert names <= set(M.EXPORTED)


def test_loads_without_gpu_and_reports_errors():
    # no GPU here: device entry points must fail with a status, not crash
    if M.device_count() > 0:
        pytest.skip("a GPU is present")
    import ctypes
    p = ctypes.c_void_p()
    assert M.lib.cuda_malloc(ctypes.byref(p), 64) != 0


def test_host_transcript_matches_oracle():
    rng = np.random.default_rng(7)
    for trial in range(30):
        label = bytes(rng.integers(0, 256, size=trial % 7, dtype=np.uint8))
        a, b = M.Transcript(label), O.Transcript(label)
        for step in range(12):
            kind = int(rng.integers(0, 5))
            lab = bytes(rng.integers(97, 123, size=int(rng.integers(0, 4)), dtype=np.uint8))
            if kind == 0:
                v = int(rng.integers(0, 1 << 63)) * 2
                a.append_u64(lab, v), b.append_u64(lab, v)
            elif kind == 1:
                v = rand_field(rng, int(rng.integers(0, 6)))
                a.append_scalars(lab, v), b.append_scalars(lab, v)
            elif kind == 2:
                v = rand_field(rng, 2)
                a.append_ext2(lab, v), b.append_ext2(lab, v)
            elif kind == 3:
                assert a.challenge_scalar(lab) == b.challenge_scalar(lab)
            else:
                assert np.array_equal(a.challenge_ext2(lab), b.challenge_ext2(lab))
        assert np.array_equal(a.state(), b.state())


def test_label_absorbed_ahead_is_the_same_transcript():
    """The sumcheck provers absorb a round's first message label before the message arrives
    (host_sumcheck.hip prove_lookahead, host_qsumcheck.hip): the label appended on its own,
    then the values appended without a label, leave both transcripts where the labelled
    append does."""
    rng = np.random.default_rng(11)
    for make in (M.Transcript, M.QTranscript):
        for lab in (b"round_message", b"x", b"abcdefgh", b"abcdefghijklmno"):
            a, b = make(b"t"), make(b"t")
            for _ in range(5):
                for _ in range(3):
                    v = rand_field(rng, 2)
                    a.append_ext2(lab, v)
                    b.append_scalars(lab, np.zeros(0, dtype=np.uint64))
                    b.append_ext2(b"", v)
                assert np.array_equal(a.challenge_ext2(b"challenge"), b.challenge_ext2(b"challenge"))
            assert np.array_equal(np.array(a.t.state, dtype=np.uint64), np.array(b.t.state, dtype=np.uint64))


def test_host_permutations_match_oracle():
    import ctypes
    rng = np.random.default_rng(8)
    g = GOLDEN["monolith_w12_permute_0_to_11"]
    s = np.array(g["input"], dtype=np.uint64)
    M.lib.mam_monolith_permute_host(s.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    assert [int(x) for x in s] == g["output"]
    for _ in range(20):
        s12 = rng.integers(0, 1 << 63, 12, dtype=np.uint64) * 2 + 1
        a = s12.copy()
        M.lib.mam_monolith_permute_host(a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        assert np.array_equal(a, O.monolith_permute(s12))
        s8 = rng.integers(0, 1 << 63, 8, dtype=np.uint64) * 2 + 1
        b = s8.copy()
        M.lib.mam_poseidon2_permute_host(b.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        assert np.array_equal(b, O.poseidon2_permute(s8, 0))
    P = (1 << 64) - (1 << 32) + 1
    for edge in ([0] * 8, [P - 1] * 8, [P] * 8, [(1 << 64) - 1] * 8, [P - 1, 0, P, 1, (1 << 64) - 1, 2, P + 5, 7]):
        s8 = np.array(edge, dtype=np.uint64)
        b = s8.copy()
        M.lib.mam_poseidon2_permute_host(b.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        assert np.array_equal(b, O.poseidon2_permute(s8, 0))


@pytest.mark.parametrize("nv,lr,nr,seed", [(4, 1, 4, 42), (6, 3, 6, 7), (10, 1, 10, 42)])
def test_host_table_matches_oracle(nv, lr, nr, seed):
    t = M.BasefoldTable.generate(nv, lr, nr, seed)
    assert np.array_equal(t.entries, O.basefold_table(nv, lr, nr, seed))


@pytest.mark.parametrize("nv,lr,nq,hash_kind", [(3, 1, 4, 0), (6, 2, 9, 0), (5, 1, 3, 1)])
def test_host_verifier_agrees_with_oracle(nv, lr, nq, hash_kind):
    rng = np.random.default_rng(50 + nv)
    ev = rand_field(rng, 1 << nv)
    cw, bh, tree = O.basefold_commit(ev, nv, lr, hash_kind)
    root = tree[-4:]
    table = M.BasefoldTable.generate(nv, lr, nv, 42)
    pt = rand_field(rng, 2 * nv)
    proof = O.basefold_open_ext2(cw, bh, root, nv, lr, hash_kind, pt, table.entries, nv, O.Transcript(b"v"), nq)
    assert M.verify_ext2(root, pt, nv, lr, hash_kind, proof, nq, table, M.Transcript(b"v"))
    bad = proof.copy()
    bad[5] = (int(bad[5]) + 1) % P
    assert not M.verify_ext2(root, pt, nv, lr, hash_kind, bad, nq, table, M.Transcript(b"v"))
    assert M.proof_len(nv, lr, nq) == O.basefold_proof_len(nv, lr, nq)
    parts = M.split_proof(proof, nv, lr, nq)
    assert len(parts["queries"]) == nq


def test_argument_checks_before_device_work():
    # rejected arguments return -2 before any device call (no GPU needed)
    import ctypes
    from morpheus_amd._lib import MatmulInstance_t
    s = MatmulInstance_t()
    s.lm, s.lk, s.ln = 3, 3, 3
    fake = ctypes.c_void_p(16)
    assert M.lib.mam_matmul_prove(ctypes.byref(s), fake, 6, 9, 0, 4, None, 0) == -2   # log_rate too large
    assert M.lib.mam_matmul_prove(ctypes.byref(s), fake, 6, 1, 2, 4, None, 0) == -2   # unknown hash
    assert M.lib.mam_matmul_prove(ctypes.byref(s), fake, 5, 1, 0, 4, None, 0) == -2   # table too small
    s.lk = 0
    assert M.lib.mam_matmul_prove(ctypes.byref(s), fake, 6, 1, 0, 4, None, 0) == -2
    assert M.lib.mam_sumcheck_proof_bincode_len(-1, 3) == -2
    buf = (ctypes.c_uint8 * 8)()
    fe = np.zeros(2, dtype=np.uint64)
    assert M.lib.mam_sumcheck_proof_bincode(None, 0, 3, fe.ctypes.data, buf, 8) == -2  # 24 bytes needed
