"""The CPU oracle and the host side of the C ABI under AddressSanitizer and
UndefinedBehaviorSanitizer (CPU only). Each suite runs in a subprocess against
a sanitized build (oracle/Makefile and open_quadtree_mapping_amd/csrc/Makefile,
target `sanitize`), with the matching compiler's runtime preloaded into the
Python interpreter; any report fails the run (UB reports abort)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _runtime(cmd):
    out = subprocess.run(cmd, capture_output=True, text=True).stdout.strip()
    return out if out and os.path.exists(out) else None


def _run(env_extra, tests):
    env = dict(os.environ)
    env.update(env_extra)
    env["ASAN_OPTIONS"] = "detect_leaks=0:detect_odr_violation=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["OMP_NUM_THREADS"] = "4"
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider"] + tests,
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    text = r.stdout + r.stderr
    assert r.returncode == 0, text[-4000:]
    assert "runtime error" not in text and "AddressSanitizer" not in text, text[-4000:]
    return text


def test_oracle_under_asan_ubsan():
    asan = _runtime(["g++", "-print-file-name=libasan.so"])
    ubsan = _runtime(["g++", "-print-file-name=libubsan.so"])
    if not (asan and ubsan):
        pytest.skip("gcc's sanitizer runtimes are not installed")
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"])
    lib = os.path.join(ROOT, "_build", "san", "libqm_oracle_san.so")
    out = _run({"LD_PRELOAD": f"{asan}:{ubsan}", "QM_ORACLE_LIB": lib},
               ["tests/test_oracle_pins.py", "tests/test_cost_arith.py"])
    print(out.strip().split("\n")[-1])


def test_host_c_abi_under_asan_ubsan():
    rt = _runtime(["/opt/rocm/bin/amdclang++", "-print-file-name=libclang_rt.asan-x86_64.so"])
    if not rt:
        import glob
        hits = glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
        rt = hits[0] if hits else None
    if not rt:
        pytest.skip("clang's ASan runtime is not installed")
    subprocess.check_call(["make", "-s", "-j8", "-C", os.path.join(ROOT, "open_quadtree_mapping_amd", "csrc"),
                           "sanitize"])
    lib = os.path.join(ROOT, "open_quadtree_mapping_amd", "_build", "san", "libqm_san.so")
    out = _run({"LD_PRELOAD": rt, "QM_LIB": lib}, ["tests/test_capi_cpu.py"])
    print(out.strip().split("\n")[-1])
