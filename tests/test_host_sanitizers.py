"""The host C++ under AddressSanitizer and UndefinedBehaviorSanitizer.

The scene builders and validation, the workload generators, the RHI
adapter's push decoding and host resources, and the CPU oracle are compiled
with -fsanitize=address,undefined (g++, host code only; the device library
they call is linked as built) together with tests/sanitize/host_driver.cpp,
which drives them with ordinary and edge-case inputs.  Any sanitizer report
(out-of-bounds access, use after free, leak, signed overflow, misaligned or
null access, invalid shift, float-to-int overflow) fails the test."""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "demont-engine_amd")
SOURCES = [
    "tests/sanitize/host_driver.cpp",
    "demont-engine_amd/csrc/host/builders.cpp",
    "demont-engine_amd/csrc/host/prep.cpp",
    "demont-engine_amd/csrc/host/scenegen.cpp",
    "demont-engine_amd/csrc/host/host_api.cpp",
    "demont-engine_amd/csrc/rhi/hip_device.cpp",
    "oracle/pt_oracle.cpp",
]
FLAGS = ["-std=c++2b", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=all", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
         "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(PKG, "csrc", "host"),
         "-I" + os.path.join(ROOT, "oracle"), "-I/opt/rocm/include"]


def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    gxx = shutil.which("g++")
    lib = os.path.join(PKG, "libdmt_hip.so")
    if not gxx or not os.path.exists(lib):
        pytest.skip("needs g++ and the built libdmt_hip.so")

    def compile_one(src):
        obj = os.path.join(tmp_path, os.path.basename(src) + ".o")
        r = subprocess.run([gxx, *FLAGS, "-c", os.path.join(ROOT, src), "-o", obj], capture_output=True, text=True)
        return src, obj, r

    with ThreadPoolExecutor(max_workers=min(7, os.cpu_count() or 2)) as ex:
        results = list(ex.map(compile_one, SOURCES))
    for src, _, r in results:
        assert r.returncode == 0, f"{src}:\n{r.stderr[-3000:]}"
    exe = os.path.join(tmp_path, "host_driver")
    link = [gxx, "-fsanitize=address,undefined", *[o for _, o, _ in results], "-L" + PKG, "-ldmt_hip",
            "-L/opt/rocm/lib", "-lamdhip64", "-lpthread", "-Wl,-rpath," + PKG, "-Wl,-rpath,/opt/rocm/lib", "-o", exe]
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0 and "asan" in r.stderr.lower() and "cannot find" in r.stderr.lower():
        pytest.skip("the sanitizer runtimes are not installed")
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:halt_on_error=1:abort_on_error=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["HIP_VISIBLE_DEVICES"] = ""   # the device's host-only paths, also on a machine with a GPU
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    report = r.stdout[-2000:] + r.stderr[-6000:]
    assert "ERROR: AddressSanitizer" not in r.stderr, report
    assert "ERROR: LeakSanitizer" not in r.stderr, report
    assert "runtime error:" not in r.stderr, report
    assert r.returncode == 0, report
    assert "host sanitizers: 0 failure(s)" in r.stdout, report
