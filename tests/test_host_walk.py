"""CPU checks of the host side's threading: the multi-threaded CU walk derives
exactly the deblocking descriptor maps, work lists and bS histogram of the
one-thread walk (tests/host_walk_check.cpp), and the worker pool runs every
index once under concurrent callers (tests/pool_check.cpp); both compiled
here from the package's host sources.  The one-thread walk itself is pinned
against the oracle by the GPU parity tests."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "oxideav-h266_amd", "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir(os.path.join(ROCM, "include", "hip")),
                    reason="needs g++ and the HIP headers")
def test_threaded_walk_matches_one_thread(tmp_path):
    exe = tmp_path / "host_walk_check"
    cmd = ["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", f"-I{CSRC}",
           os.path.join(ROOT, "tests", "host_walk_check.cpp"), os.path.join(CSRC, "vvclf_host.cpp"),
           os.path.join(CSRC, "synth.cpp"), f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib", "-lpthread",
           "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("identical") == 9


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir(os.path.join(ROCM, "include", "hip")),
                    reason="needs g++ and the HIP headers")
def test_worker_pool_runs_every_index_once(tmp_path):
    """The host worker pool behind the threaded walk and the staging copies,
    with six threads calling it at once."""
    exe = tmp_path / "pool_check"
    cmd = ["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", f"-I{CSRC}",
           os.path.join(ROOT, "tests", "pool_check.cpp"), os.path.join(CSRC, "vvclf_host.cpp"), f"-L{ROCM}/lib",
           "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib", "-lpthread", "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "missed or repeated indices: 0" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir(os.path.join(ROCM, "include", "hip")),
                    reason="needs g++ and the HIP headers")
def test_lmcs_table_matches_oracle_per_sample(tmp_path):
    """The host's LMCS inverse table, built by one walk over the samples,
    equals the oracle's per-sample bin scan and inverse map for every u16
    input (tests/lmcs_lut_check.cpp)."""
    exe = tmp_path / "lmcs_lut_check"
    orc = os.path.join(ROOT, "oracle")
    obj = tmp_path / "vvclf_oracle.o"
    subprocess.run(["gcc", "-O2", "-c", "-o", str(obj), os.path.join(orc, "vvclf_oracle.c")], check=True,
                   capture_output=True, timeout=300)
    cmd = ["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", f"-I{CSRC}", f"-I{orc}",
           os.path.join(ROOT, "tests", "lmcs_lut_check.cpp"), os.path.join(CSRC, "vvclf_host.cpp"), str(obj),
           f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib", "-lpthread", "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "lmcs tables identical" in r.stdout
