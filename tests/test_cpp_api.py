"""The C++ host API (include/gpu_bfs.hpp) through its compiled test program,
tests/cpp/test_api.cpp (built by `make -C tests/cpp`, which build() runs).
Its host half needs no GPU; its GPU half checks every method against closed
forms (grid Manhattan distances, path |i-j|, star 0/1/2, component
sentinels) and the reference's error texts."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "test_api")


def run(*args, timeout):
    if not os.path.exists(BIN):
        raise AssertionError(f"{BIN} is not built: run `make -C tests/cpp`")
    print("running", BIN, *args, flush=True)
    r = subprocess.run([BIN, *args], capture_output=True, text=True, timeout=timeout)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK: 0 failed checks" in r.stdout


def test_cpp_api_host_side():
    run("--cpu", timeout=60)


@pytest.mark.gpu
def test_cpp_api_on_gpu():
    run(timeout=100)


SAN_BIN = os.path.join(HERE, "cpp", "test_sanitize")


def test_host_code_under_sanitizers():
    """tests/cpp/test_sanitize.cpp: the library's host sources (graph builders,
    generators, Matrix Market parser) and the oracle built with ASan + UBSan
    (+ LeakSanitizer) by gcc and driven over empty / ragged / truncated /
    mutated inputs; any sanitizer report fails the run."""
    if not os.path.exists(SAN_BIN):
        raise AssertionError(f"{SAN_BIN} is not built: run `make -C tests/cpp`")
    env = dict(os.environ, ASAN_OPTIONS="halt_on_error=1:detect_leaks=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([SAN_BIN], capture_output=True, text=True, timeout=120, env=env)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK:" in r.stdout and ", 0 failed" in r.stdout
