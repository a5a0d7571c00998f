"""Build the oracle (test infrastructure): oracle/liboracle.so from sdesim_oracle.cpp.

The reference itself is Rust and cannot be built in this environment (no cargo/rustc, and its
crates are not vendored), so oracle/_ref stays empty; the oracle is a C++ restatement.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "liboracle.so")


def build():
    src = os.path.join(HERE, "sdesim_oracle.cpp")
    tmp = OUT + ".tmp"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", tmp, src, "-lpthread"], check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
