"""Builds the native libraries of the package in-tree.

libvvclf.so        HIP kernels (gfx950) + the C ABI of include/vvclf.h
libvvclf_synth.so  seeded synthetic workload generator (host C++)

Each source is compiled to its own object in parallel, then linked.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
ARCH = os.environ.get("VVCLF_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["vvclf_plain.hip", "vvclf_fused.hip", "vvclf_derive.hip", "vvclf_host.cpp", "vvclf_api.cpp",
               "vvclf_batch.cpp", "vvclf_server.cpp", "vvclf_noalloc.cpp", "vvclf_recon.hip",
               "vvclf_syntax.cpp", "vvclf_enc.hip"]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             # packed FP32 instructions are disabled for all device code
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-value"]


def _run(cmd):
    r = subprocess.run(cmd, cwd=HERE, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    _run(["hipcc", *HIP_FLAGS, *lang, "-c", "-o", obj, os.path.join(CSRC, src)])
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    out = os.path.join(HERE, "libvvclf.so")
    _run(["hipcc", "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", out + ".tmp", *objs, "-lpthread"])
    os.replace(out + ".tmp", out)
    synth = os.path.join(HERE, "libvvclf_synth.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", synth + ".tmp", os.path.join(CSRC, "synth.cpp")])
    os.replace(synth + ".tmp", synth)
    if verbose:
        print("built", out, synth)
    return out


if __name__ == "__main__":
    build(verbose=True)
