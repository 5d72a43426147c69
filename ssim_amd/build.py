"""Build libssim_hip.so for gfx950 with hipcc (in-tree, next to this file)."""
from __future__ import annotations

import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libssim_hip.so")
SOURCES = ["kernels.hip", "batch_kernels.hip", "context.hip", "topology.hip", "solvers.hip", "batch.hip", "noalloc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed-FP32 instructions stay off (they are generated from ordinary float code otherwise)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"),
]


# kernels that must not spill registers to scratch (the PCG iteration kernels run thousands of times per step)
NO_SCRATCH = re.compile(r"pcg_(spmv|update)")


def _resource_usage(stderr: str):
    """{kernel symbol: (VGPRs, scratch bytes per lane)} from -Rpass-analysis=kernel-resource-usage remarks."""
    out, name, vgpr = {}, None, None
    for line in stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            continue
        m = re.search(r"remark:\s+VGPRs: (\d+)", line)
        if m and name:
            vgpr = int(m.group(1))
            continue
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and name:
            out[name] = (vgpr, int(m.group(1)))
            name = None
    return out


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-Rpass-analysis=kernel-resource-usage", *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stderr=subprocess.PIPE, text=True)
    if res.returncode != 0:
        sys.stderr.write("\n".join(l for l in res.stderr.splitlines() if "remark:" not in l) + "\n")
        raise subprocess.CalledProcessError(res.returncode, cmd)
    usage = _resource_usage(res.stderr)
    spills = {k: v for k, v in usage.items() if NO_SCRATCH.search(k) and v[1] > 0}
    if spills:
        os.remove(tmp)
        raise RuntimeError(f"PCG kernels spill to scratch (VGPRs, bytes/lane): {spills}")
    if verbose:
        for k, (vg, sc) in sorted(usage.items()):
            if sc:
                print(f"  scratch {sc:5d} B/lane  VGPRs {vg:3d}  {k}", flush=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
