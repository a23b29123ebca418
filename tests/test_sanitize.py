"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

Compiles tests/sanitize/driver.cpp with the product's plan code
(msmpb_amd/csrc/msm_plan.cpp) and both oracles using gcc/g++
-fsanitize=address,undefined, then runs it: plan building and rebuilding
for every order, level count and method (rejected inputs included), and
oracle evaluations.  Any sanitizer report fails the test.
"""
import os
import shutil
import subprocess

import pytest

from helpers import ROOT

SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g",
       "-O1"]


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_host_code_clean_under_sanitizers(tmp_path):
    objs = []
    for src in ("oracle/msm_oracle.c", "oracle/pme_oracle.c"):
        o = tmp_path / (os.path.basename(src) + ".o")
        subprocess.check_call(["gcc", "-std=c11", *SAN, "-c", os.path.join(ROOT, src), "-o", str(o)])
        objs.append(str(o))
    exe = tmp_path / "driver"
    subprocess.check_call(["g++", "-std=c++17", *SAN, os.path.join(ROOT, "tests", "sanitize", "driver.cpp"),
                           os.path.join(ROOT, "msmpb_amd", "csrc", "msm_plan.cpp"), *objs, "-lm",
                           "-o", str(exe)])
    # verify_asan_link_order=0: the environment may preload a library ahead
    # of the sanitizer runtime
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
    assert r.stdout.startswith("ok built=")
