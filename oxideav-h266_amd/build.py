"""Builds the native libraries of the package in-tree.

libvvclf.so        HIP kernels (gfx950) + the C ABI of include/vvclf.h
libvvclf_synth.so  seeded synthetic workload generator (host C++)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ARCH = os.environ.get("VVCLF_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["vvclf_plain.hip", "vvclf_fused.hip", "vvclf_derive.hip", "vvclf_host.cpp", "vvclf_api.cpp", "vvclf_batch.cpp",
               "vvclf_server.cpp"]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
             # packed FP32 instructions are disabled for all device code
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-value"]


def _run(cmd):
    r = subprocess.run(cmd, cwd=HERE, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    out = os.path.join(HERE, "libvvclf.so")
    _run(["hipcc", *HIP_FLAGS, "-o", out + ".tmp", *srcs, "-lpthread"])
    os.replace(out + ".tmp", out)
    synth = os.path.join(HERE, "libvvclf_synth.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", synth + ".tmp", os.path.join(CSRC, "synth.cpp")])
    os.replace(synth + ".tmp", synth)
    if verbose:
        print("built", out, synth)
    return out


if __name__ == "__main__":
    build(verbose=True)
