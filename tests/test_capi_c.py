"""The C ABI from C: tests/capi_check.c is compiled with gcc as C99 against
include/d2hip.h and linked with libd2hip.so, so header/library drift (a
signature or a symbol that no longer matches) fails the build.  On the CPU it
only resolves the entry points; on the GPU it runs the reference fixtures
through the ABI (no Python in between)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "delta2_amd")


def _build(tmp_path):
    exe = str(tmp_path / "capi_check")
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "capi_check.c"), "-L", LIBDIR, "-ld2hip", "-lm", f"-Wl,-rpath,{LIBDIR}",
           "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return exe


def test_c_program_builds_and_links(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe, "link"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "entry points linked" in r.stdout, r.stdout + r.stderr


@pytest.mark.gpu
def test_c_program_runs_the_reference_fixtures(tmp_path):
    exe = _build(tmp_path)
    print("[capi_check] running the C program", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "capi_check: ok" in r.stdout, r.stdout + r.stderr
