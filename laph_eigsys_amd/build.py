"""Builds liblapheig.so in-tree for gfx950 (make in laph_eigsys_amd/csrc)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs=8, clean=False):
    csrc = os.path.join(HERE, "csrc")
    if clean:
        subprocess.run(["make", "-C", csrc, "clean"], check=True, stdout=subprocess.DEVNULL)
    subprocess.run(["make", "-C", csrc, f"-j{jobs}"], check=True)
    return os.path.join(HERE, "liblapheig.so")


if __name__ == "__main__":
    build(clean="--clean" in sys.argv)
