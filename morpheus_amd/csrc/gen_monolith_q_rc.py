"""Writes monolith_q_rc.h: the round constants of Monolith-w12 over q = 2^64 - 2^32 - 31 as the
zk-torch-4 transcript derives them (zk-torch-4/src/crypto/monolith.rs:67-110): for rounds 1..5
and positions 0..11, the first 8 bytes (little-endian) of
SHA-256("AGL-Monolith-RC-v1" || round u32 LE || pos u32 LE || salt u32 LE), taking the first
salt = 0, 1, ... whose value is below q. Rounds 0 and 6 have zero constants.

Run from this directory: python gen_monolith_q_rc.py"""
import hashlib
import struct

Q = 0xFFFFFFFEFFFFFFE1


def constant(rnd, pos):
    salt = 0
    while True:
        d = hashlib.sha256(b"AGL-Monolith-RC-v1" + struct.pack("<III", rnd, pos, salt)).digest()
        v = int.from_bytes(d[:8], "little")
        if v < Q:
            return v
        salt += 1


def main():
    rows = [[constant(r, i) for i in range(12)] for r in range(1, 6)]
    with open("monolith_q_rc.h", "w") as f:
        f.write("// Generated by gen_monolith_q_rc.py: Monolith-w12 round constants over q for rounds 1..5\n"
                "// (zk-torch-4/src/crypto/monolith.rs:67-110). Do not edit.\n#pragma once\n\n"
                "#define MONOQ_RC_ROWS \\\n" + " \\\n".join(
                    ("    {" + ", ".join(f"0x{v:016x}ULL" for v in row) + "}" + ("," if k < 4 else ""))
                    for k, row in enumerate(rows)) + "\n")


if __name__ == "__main__":
    main()
