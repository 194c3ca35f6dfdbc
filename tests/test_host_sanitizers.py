"""The host parsers and scene code (JSON scene files, OBJ meshes, PNG and
Radiance images, procedural scenes, the .hdr writer) and the oracle, built
with AddressSanitizer and UndefinedBehaviorSanitizer (the Makefiles' host-san
and san targets, plain g++), run over valid and damaged inputs in a child
process with the sanitizer runtimes preloaded.  Any report fails the test.
Host only."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _runtime(name):
    r = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True)
    p = r.stdout.strip()
    return p if r.returncode == 0 and os.path.isabs(p) and os.path.exists(p) else None


@pytest.fixture(scope="module")
def san_libs():
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("no sanitizer runtimes for g++ on this host")
    subprocess.run(["make", "-C", os.path.join(ROOT, "pico_amd", "csrc"), "host-san"], check=True, capture_output=True)
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "san"], check=True, capture_output=True)
    return asan, ubsan


def test_parsers_and_oracle_under_sanitizers(san_libs):
    asan, ubsan = san_libs
    env = dict(os.environ, LD_PRELOAD=f"{asan}:{ubsan}",
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", PICO_SAN_MUTATIONS="600")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "host_san_driver.py")], env=env,
                       capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert r.stdout.strip().endswith("done"), r.stdout
