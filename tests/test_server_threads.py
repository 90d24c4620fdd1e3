"""The batching server's threading and admission, on the host: server.cpp is
built against a stand-in prover library (tests/server_standin.cpp) and run
under ThreadSanitizer.  Checks: no data race or lock misuse; every job's
result is the one its inputs give, alone or in a mix; a memory budget that
fits k contexts admits exactly k jobs at once and is never exceeded; idle
workers release contexts of other sizes so the queue moves; a job whose
creation fails reports its error text in the job; destroying the server
finishes every submitted job (and a Keccak configuration is refused); jobs submitted together form groups of one
shape up to the group limit; with the self-check on, no proof is published
unverified and a rejected one fails only its own job; repeated batches reuse
the workers' contexts (size affinity) instead of creating new ones."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "whir-p3_amd", "csrc")
SRCS = [os.path.join(CSRC, "server.cpp"), os.path.join(ROOT, "tests", "server_standin.cpp")]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


def build(tmp_path, sanitize):
    out = str(tmp_path / ("server_tsan" if sanitize else "server_plain"))
    flags = ["-O1", "-g", "-fsanitize=thread"] if sanitize else ["-O2"]
    subprocess.run(["g++", "-std=c++17", "-Wall", *flags, "-I", CSRC, *SRCS, "-o", out, "-lpthread"], check=True,
                   capture_output=True, text=True)
    return out


@pytest.fixture(scope="module")
def binaries(tmp_path_factory):
    d = tmp_path_factory.mktemp("server")
    return build(d, True), build(d, False)


@pytest.mark.parametrize("scenario", ["mix", "budget", "trim", "oversized", "drain", "groups", "self-check", "affinity"])
def test_server_under_thread_sanitizer(binaries, scenario):
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    r = subprocess.run([binaries[0], scenario], capture_output=True, text=True, timeout=240, env=env)
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert " ok" in r.stdout


def test_server_timed_waits(binaries):
    # timed waits go through pthread_cond_clockwait, which the toolchain's
    # ThreadSanitizer does not intercept: this scenario runs in a plain build
    r = subprocess.run([binaries[1], "mix-timed"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
