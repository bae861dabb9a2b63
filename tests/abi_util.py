"""Build and location of tests/abi_harness.c, a plain C client of the ABI
(gcc -std=c99 against include/ugx.h, linked with ugacc_amd/libugx.so)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "abi_harness.c")
BIN = os.path.join(ROOT, "tests", "_bin", "abi_harness")


def build():
    """Compiles the harness (a few hundred ms); needs libugx.so built."""
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    cmd = ["gcc", "-std=c99", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"), SRC,
           "-L", os.path.join(ROOT, "ugacc_amd"), "-lugx", "-Wl,-rpath,$ORIGIN/../../ugacc_amd", "-o", BIN + ".tmp"]
    subprocess.run(cmd, check=True)
    os.replace(BIN + ".tmp", BIN)
    return BIN


def harness():
    """Path of the built harness (built first when missing)."""
    if not os.path.exists(BIN):
        build()
    return BIN
