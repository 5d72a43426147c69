"""Builds the C++ drivers of the GPU tests against the in-tree libssim_hip.so:
tests/cpp/build/driver (g++, the front end include/ssim_hip.hpp) and tests/cpp/build/time_step_driver (hipcc,
ssim::fem::basic_time_step and the mathprim subset under include/ssim and include/mathprim)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "build", "driver")
OUT_TS = os.path.join(HERE, "build", "time_step_driver")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    lib = ["-L", os.path.join(ROOT, "ssim_amd"), "-lssim_hip", "-Wl,-rpath,$ORIGIN/../../../ssim_amd"]
    cmds = [
        ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "driver.cpp"),
         *lib, "-o", OUT],
        [HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-Wall", "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "time_step_driver.hip"),
         *lib, "-o", OUT_TS],
    ]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
