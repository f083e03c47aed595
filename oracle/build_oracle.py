"""Builds the CPU oracle (test infrastructure) into oracle/_build/.

    python oracle/build_oracle.py

-ffp-contract=off keeps every multiply and add separately rounded except
the explicit fma()/fmaf() calls, which is the op sequence the kernels use.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build():
    out_dir = os.path.join(HERE, "_build")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "libbm3d_oracle.so")
    tmp = out + ".tmp"
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-mfma",
           "-Wall", "-Wextra", "-o", tmp, os.path.join(HERE, "bm3d_oracle.c"), "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build())
    sys.exit(0)
