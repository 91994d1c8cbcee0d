"""The oracle and the host-side TOML reader under AddressSanitizer and
UndefinedBehaviorSanitizer (the reference's debug build flags, cmake/gnu-debug.cmake:25,28):
the oracle's own pin tests run against the sanitized oracle library in a child process,
and the TOML reader parses the example job files and refuses malformed texts without an
out-of-bounds access.  CPU only."""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _asan_runtime():
    out = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = out.stdout.strip()
    if out.returncode != 0 or not os.path.isabs(path) or not os.path.exists(path):
        pytest.skip("no libasan runtime next to gcc")
    return path


def _env(extra):
    env = dict(os.environ)
    env.update(ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:exitcode=86",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=87")
    env.update(extra)
    return env


def test_toml_reader_under_sanitizers(tmp_path):
    exe = str(tmp_path / "san_toml")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-o", exe, os.path.join(ROOT, "tests", "san_toml.cpp")],
                   check=True)
    files = sorted(glob.glob(os.path.join(ROOT, "icarat-contact_amd", "host", "examples", "*.toml")))
    res = subprocess.run([exe] + files, capture_output=True, text=True, env=_env({}), timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr


def test_oracle_pins_under_sanitizers():
    asan = _asan_runtime()
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True)
    lib = os.path.join(ROOT, "oracle", "_build_san", "libicx_oracle.so")
    tests = sorted(glob.glob(os.path.join(ROOT, "tests", "test_oracle_*.py"))) + [
        os.path.join(ROOT, "tests", "test_golden.py")]
    env = _env({"LD_PRELOAD": asan, "ICX_ORACLE_LIB": lib, "OMP_NUM_THREADS": "4"})
    res = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu"] + tests,
                         capture_output=True, text=True, env=env, cwd=ROOT, timeout=900)
    assert res.returncode == 0, (res.stdout[-3000:] + res.stderr[-3000:])
    assert "passed" in res.stdout
