"""The ANSI-C host layer (case I/O, LCG and block draw, initial conditions,
backup files, network parameter files, shear-layer statistics) built with
gcc's AddressSanitizer and UndefinedBehaviorSanitizer and driven over four
meshes by tests/host_sanitize.c: no report, every mesh completes.  Host code
only (no device call)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_host_layer_under_asan_and_ubsan(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    exe = str(tmp_path / "host_sanitize")
    src = [os.path.join(ROOT, "freeshearlayer_amd", "host", f) for f in ("fsl_host.c", "fsl_stats.c")]
    cmd = [gcc, "-std=c99", "-g", "-O1", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-I" + os.path.join(ROOT, "include")] + src + \
          [os.path.join(HERE, "host_sanitize.c"), "-lm", "-o", exe]
    b = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    if b.returncode != 0 and "asan" in b.stdout.lower():
        pytest.skip("sanitizer runtime not available: " + b.stdout[-200:])
    assert b.returncode == 0, b.stdout
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, str(tmp_path)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "runtime error" not in r.stdout and "ERROR: AddressSanitizer" not in r.stdout, r.stdout[-4000:]
    assert r.stdout.count(" ok (") == 4, r.stdout
