"""Regenerates the data-only entries of reference_vectors.json from the reference's
source text (run in the build container, where the reference checkout is mounted).

It reads constants and known-answer numbers as text; nothing of the reference is
imported or executed. Tests only read the JSON.
"""
import json
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = __file__.replace("transcribe_vectors.py", "reference_vectors.json")


def grab(src, name):
    i = src.index("const " + name)
    k = src.index("=", i)
    j = src.index("];", k)
    return [int(x, 16) for x in re.findall(r"0x[0-9a-fA-F]+", src[k:j])]


def main():
    d = json.load(open(OUT))
    src = open(f"{REF}/zk-torch-3/src/transcript.rs").read()
    ext, inn, diag = grab(src, "RC_EXT:"), grab(src, "RC_INT:"), grab(src, "DIAG_8:")
    assert (len(ext), len(inn), len(diag)) == (64, 22, 8)
    d["poseidon2_w8_constants"] = {
        "source": "zk-torch-3/src/transcript.rs:14-63 (RC_EXT, RC_INT, DIAG_8)",
        "rc_ext": [ext[8 * r:8 * r + 8] for r in range(8)], "rc_int": inn, "diag": diag}
    mono = open(f"{REF}/goldilocks-cuda-rs/src/cpu_monolith.rs").read()
    i = mono.index("let expected: [u64; 12]")
    nums = [int(x) for x in re.findall(r"\d{6,}", mono[i:mono.index("];", i)])]
    assert nums == d["monolith_w12_permute_0_to_11"]["output"]
    json.dump(d, open(OUT, "w"), indent=1)


if __name__ == "__main__":
    main()
