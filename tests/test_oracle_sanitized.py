"""The oracle's own CPU tests, rerun against an AddressSanitizer +
UndefinedBehaviorSanitizer build of oracle/xdg_oracle.c (make -C oracle san)
in a child process with the ASan runtime preloaded: an out-of-bounds
access, use after free or undefined operation in the oracle fails them.
Host code only."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "oracle", "_san", "liboracle.so")


def _runtime(name):
    r = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True)
    p = r.stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_oracle_suite_under_asan_ubsan():
    asan = _runtime("libasan.so")
    if asan is None:
        pytest.skip("no ASan runtime")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True)
    env = dict(os.environ)
    env.update(LD_PRELOAD=asan, XDG_ORACLE_LIB=SAN,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="4")
    probe = subprocess.run([sys.executable, "-c", "import sys; sys.path.insert(0, 'oracle'); import pyoracle as O; "
                            "print(O.lib()._name); print('libasan' in open('/proc/self/maps').read())"],
                           cwd=ROOT, env=env, capture_output=True, text=True, timeout=120)
    assert probe.stdout.split() == [SAN, "True"], probe.stdout + probe.stderr
    tests = ["tests/test_oracle_core.py", "tests/test_oracle_golden.py", "tests/test_oracle_find_element.py",
             "tests/test_oracle_mesh_walk.py", "tests/test_oracle_transport.py", "tests/test_oracle_adjoint.py",
             "tests/test_overlap_oracle.py"]
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", *tests],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error:" not in r.stderr, tail
