"""Build step for libmorpheus_amd.so (hipcc, gfx950) via morpheus_amd/csrc/Makefile."""
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))


def build(force=True, jobs=8):
    cmd = ["make", "-C", os.path.join(_HERE, "csrc"), f"-j{jobs}"]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)
    return os.path.join(_HERE, "libmorpheus_amd.so")
