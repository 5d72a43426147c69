"""Builds tests/cpp/build/driver (the C++ front-end driver) against the in-tree libssim_hip.so."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "build", "driver")


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "driver.cpp"), "-L", os.path.join(ROOT, "ssim_amd"), "-lssim_hip",
           "-Wl,-rpath,$ORIGIN/../../../ssim_amd", "-o", OUT]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
