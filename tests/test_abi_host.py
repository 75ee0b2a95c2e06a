"""The C ABI's host logic on the CPU, under AddressSanitizer.

zpc_amd/csrc/abi.hip is compiled host-only and linked against a mock of the
HIP runtime (tests/host_mock/mock_hip.cpp: host memory, streams, events and
graphs as checked handles, stream capture with CUDA's join rules) and no-op
stand-ins for the kernels' translation units (tests/host_mock/stubs.cpp). The
driver (tests/host_mock/driver.cpp) runs the policy life cycle, the sorted
driver eagerly and from a captured graph, the closure selection, the batched
scenes (scene table validation and upload, capture, autotuning), the layout
adapters' validation, a checkpoint round trip (io.hip, compiled host-only as
well: its files are real, its device memory is the mock's host memory) with a
corrupted and a mismatched load, the bgeo writer, and teardown; every
call must return the documented code, and AddressSanitizer must report no
invalid access and no leak. No kernel runs: this checks the host side only.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"
MOCK = os.path.join(ROOT, "tests", "host_mock")
CSRC = os.path.join(ROOT, "zpc_amd", "csrc")
INC = os.path.join(ROOT, "include")

def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    assert r.returncode == 0, " ".join(cmd) + "\n" + r.stdout[-2000:] + r.stderr[-4000:]
    return r


def _build(tmp) -> str:
    src = open(os.path.join(CSRC, "abi.hip")).read()
    src = src.replace('"../../include/zsmpm.h"', '"zsmpm.h"')
    abi = os.path.join(tmp, "abi_host.hip")
    open(abi, "w").write(src)
    host = [HIPCC, "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "--cuda-host-only", "-fPIC",
            "-Xarch_host", "-fsanitize=address", "-fno-omit-frame-pointer", "-Wno-unused-command-line-argument",
            f"-I{INC}", f"-I{CSRC}"]
    _run(host + ["-c", abi, "-o", "abi.o"], tmp)
    io_src = open(os.path.join(CSRC, "io.hip")).read().replace('"../../include/zsmpm.h"', '"zsmpm.h"')
    io = os.path.join(tmp, "io_host.hip")
    open(io, "w").write(io_src)
    _run(host + ["-c", io, "-o", "io.o"], tmp)
    _run(host + ["-x", "hip", "-c", os.path.join(MOCK, "stubs.cpp"), "-o", "stubs.o"], tmp)
    # the host-only object references its (absent) device code object by name
    syms = _run(["nm", "-u", "abi.o", "io.o"], tmp).stdout
    fat = sorted(set(re.findall(r"__hip_fatbin_[0-9a-f]+", syms)))
    open(os.path.join(tmp, "fatbin.c"), "w").write("".join(f"char {s}[64];\n" for s in fat))
    gxx = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address", "-fno-omit-frame-pointer"]
    _run(["gcc", "-c", "fatbin.c", "-o", "fatbin.o"], tmp)
    _run(gxx + ["-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-c", os.path.join(MOCK, "mock_hip.cpp"),
                "-o", "mock.o"], tmp)
    _run(gxx + [f"-I{INC}", "-c", os.path.join(MOCK, "driver.cpp"), "-o", "driver.o"], tmp)
    exe = os.path.join(tmp, "abi_host_probe")
    _run(gxx + ["driver.o", "abi.o", "io.o", "stubs.o", "mock.o", "fatbin.o", "-o", exe], tmp)
    return exe


@pytest.mark.skipif(not os.path.exists(HIPCC) or shutil.which("g++") is None, reason="needs hipcc and g++")
def test_abi_host_logic_under_asan(tmp_path):
    exe = _build(str(tmp_path))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0 and "PASSED" in r.stdout, r.stdout[-3000:] + r.stderr[-6000:]
