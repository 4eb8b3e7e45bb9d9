"""The oracle and the host side of the C ABI under AddressSanitizer and
UndefinedBehaviorSanitizer (CPU only).

The sanitizer builds (oracle/Makefile `san`: g++; nt2_amd/csrc/Makefile
`san`: hipcc with -fsanitize after -Xarch_host, so only host code is
instrumented) are loaded into a Python started with the matching runtime
preloaded, and run tests/test_oracle.py (minus the exhaustive walks over
every float, which repeat the same code paths 2^31 times) and
tests/test_abi.py.  A probe library with a planted heap overflow shows that
this set-up reports errors."""
import os
import subprocess
import sys
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GCC_ASAN = "/usr/lib/gcc/x86_64-linux-gnu/11/libasan.so"
CLANG_ASAN = "/opt/rocm/lib/llvm/lib/clang/22/lib/linux/libclang_rt.asan-x86_64.so"


def _env(preload, **extra):
    env = dict(os.environ)
    env["LD_PRELOAD"] = preload
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0:exitcode=23"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1:exitcode=24"
    env.update(extra)
    return env


def _need(path):
    if not os.path.exists(path):
        pytest.fail(f"{path} is missing from this image")


def test_planted_overflow_is_reported(tmp_path):
    _need(GCC_ASAN)
    src = tmp_path / "probe.cpp"
    src.write_text(textwrap.dedent("""
        #include <cstdlib>
        extern "C" int probe(int n) { int* p = (int*)std::malloc(n * sizeof(int)); int v = p[n]; std::free(p); return v; }
    """))
    lib = tmp_path / "libprobe.so"
    subprocess.run(["g++", "-O0", "-g", "-fPIC", "-shared", "-fsanitize=address", "-o", str(lib), str(src)], check=True)
    code = f"import ctypes; ctypes.CDLL({str(lib)!r}).probe(4)"
    r = subprocess.run([sys.executable, "-c", code], env=_env(GCC_ASAN), capture_output=True, text=True, timeout=120)
    assert r.returncode == 23 and "heap-buffer-overflow" in r.stderr, r.stderr[-2000:]


def test_oracle_under_asan_ubsan():
    _need(GCC_ASAN)
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True)
    lib = os.path.join(ROOT, "oracle", "_san", "libnt2ref.so")
    env = _env(GCC_ASAN, NT2REF_LIB=lib)
    # the sanitizer build is the one loaded, with the runtime in the process
    chk = ("import sys; sys.path.insert(0, 'oracle'); import nt2ref; nt2ref.lib(); "
           "m = open('/proc/self/maps').read(); assert nt2ref.LIB_PATH.endswith('_san/libnt2ref.so'); "
           "assert '_san/libnt2ref.so' in m and 'libasan' in m; print('ok')")
    r = subprocess.run([sys.executable, "-c", chk], cwd=ROOT, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ok" in r.stdout, r.stderr[-2000:]
    r = subprocess.run([sys.executable, "-m", "pytest", "tests/test_oracle.py", "-q", "-p", "no:cacheprovider",
                        "-k", "not exhaustive and not log_histogram"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr


def test_abi_host_code_under_asan_ubsan():
    _need(CLANG_ASAN)
    subprocess.run(["make", "-s", "-j8", "-C", os.path.join(ROOT, "nt2_amd", "csrc"), "san"], check=True,
                   capture_output=True)
    lib = os.path.join(ROOT, "nt2_amd", "lib", "san", "libnt2hip.so")
    env = _env(CLANG_ASAN, NT2H_LIB=lib)
    chk = ("import nt2_amd; nt2_amd.lib(); m = open('/proc/self/maps').read(); "
           "assert 'lib/san/libnt2hip.so' in m and 'clang_rt.asan' in m; print('ok')")
    r = subprocess.run([sys.executable, "-c", chk], cwd=ROOT, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ok" in r.stdout, r.stderr[-2000:]
    r = subprocess.run([sys.executable, "-m", "pytest", "tests/test_abi.py", "-q", "-p", "no:cacheprovider"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
