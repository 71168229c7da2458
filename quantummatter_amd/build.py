"""Builds libqmhip.so for gfx950 in-tree (quantummatter_amd/libqmhip.so)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int = 8) -> str:
    csrc = os.path.join(HERE, "csrc")
    # Always recompile: the object files are rebuilt from the current sources.
    subprocess.run(["make", "-s", "-C", csrc, "clean"], check=True)
    subprocess.run(["make", "-s", f"-j{jobs}", "-C", csrc], check=True)
    return os.path.join(HERE, "libqmhip.so")


if __name__ == "__main__":
    print(build(int(sys.argv[1]) if len(sys.argv) > 1 else 8))
