"""Compile the CPU oracle (test infrastructure) to oracle/lib/libelmm_oracle.so.

gcc -O2 -fopenmp -ffp-contract=off -fno-fast-math: every floating-point
operation of the restatement is rounded as written.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "libelmm_oracle.so")
FLAGS = ["-std=gnu11", "-O2", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
         "-Wall", "-Wextra", "-Wno-unused-parameter"]


def build():
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = ["gcc"] + FLAGS + ["-o", tmp, os.path.join(HERE, "elmm_oracle.c"), "-lm"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
