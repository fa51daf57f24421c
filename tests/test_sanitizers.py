"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU): the
volume builder, the .vbx writer / readers and file-set loader of the product
library, and the CPU oracle (two frames of every stage, every tracking mode),
built with g++ -fsanitize=address,undefined by tests/sanitize/Makefile and run as
one process; any report fails the test."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "sanitize")], check=True, capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([os.path.join(ROOT, "tests", "sanitize", "_build", "sanitize_driver"), str(tmp_path)],
                       capture_output=True, text=True, env=env, timeout=600)
    report = r.stdout + r.stderr
    assert r.returncode == 0, report[-4000:]
    assert "runtime error" not in report and "AddressSanitizer" not in report, report[-4000:]
    assert "sanitize driver: ok" in r.stdout
