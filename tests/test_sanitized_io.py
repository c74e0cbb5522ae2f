"""The drop-in's file side (libpgio: config parsing, all.h5 loading with the
config overrides, resume, the output writers) under the address and
undefined-behaviour sanitizers: libpgio is rebuilt with them and every test
of tests/test_dropin_io.py runs against that build in a child process with
the sanitizer runtime preloaded.  Any report fails the test.  CPU only."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _runtime(name):
    try:
        p = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True)
    except OSError:
        return None
    path = p.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_file_side_is_clean_under_address_and_ub_sanitizers(tmp_path):
    asan = _runtime("libasan.so")
    if asan is None:
        pytest.skip("the compiler's address-sanitizer runtime is not installed")
    from perigrain_amd import build
    lib = build.build_io_sanitized(str(tmp_path))
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": asan,
        # the interpreter itself is not instrumented and keeps memory until exit
        "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=0",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
        "PERIGRAIN_LIBPGIO": lib,
    })
    p = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        "-m", "not gpu", os.path.join(ROOT, "tests", "test_dropin_io.py")],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = p.stdout + p.stderr
    print(out[-3000:])
    assert "AddressSanitizer" not in out and "runtime error:" not in out, out[-3000:]
    assert p.returncode == 0, out[-3000:]
    assert " passed" in p.stdout
