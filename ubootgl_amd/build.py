"""Build step: compiles libubg.so for gfx950 (hipcc cross-compiles without a
GPU) and the C++ drop-in example, in-tree."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def build_library(jobs=8, verbose=False):
    cmd = ["make", "-B", f"-j{jobs}", "-C", os.path.join(HERE, "csrc")]
    out = subprocess.run(cmd, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("libubg.so build failed:\n" + out.stdout[-4000:] + out.stderr[-8000:])
    if verbose:
        print(out.stdout)
    return os.path.join(HERE, "libubg.so")


if __name__ == "__main__":
    print(build_library(verbose=True))
