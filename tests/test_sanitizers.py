"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer.

Two drivers are compiled here with `-fsanitize=address,undefined
-fno-sanitize-recover=all` (host code only; no GPU code is built):

- tests/native/sanitize_host.cpp links the library's host C++ -- the .asc
  reader and schedule-to-planes builder (which parses untrusted bytes), the
  stone-width solver, the camera and cut builders -- and feeds it the four
  catalogue designs, hand-written malformed inputs and 1 600 seeded
  mutations of the catalogue bytes, calling every accessor on what parses;
- tests/native/sanitize_oracle.c links the C oracle and runs small frames
  through every transport arm and lighting model, a metric image with its
  per-pixel scintillation maps, the fan-grid pose metrics, the stone width
  and the .asc plane builder.

Any sanitizer report aborts the driver, so a zero exit is the pass.  CPU only.
"""
import glob
import json
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "indicatrix_amd", "csrc")
SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                                reason="needs gcc / g++")


def _run(exe, args):
    p = subprocess.run([exe] + args, capture_output=True, text=True, env=ENV, timeout=600)
    assert p.returncode == 0, p.stderr[-4000:]
    assert "runtime error" not in p.stderr and "AddressSanitizer" not in p.stderr, p.stderr[-4000:]
    return json.loads(p.stdout.strip().splitlines()[-1])


def test_host_cpp_under_sanitizers(tmp_path):
    exe = str(tmp_path / "sanitize_host")
    subprocess.run(["g++", "-std=gnu++17", *SAN, "-o", exe, os.path.join(ROOT, "tests", "native", "sanitize_host.cpp"),
                    os.path.join(CSRC, "asc_design.cpp"), os.path.join(CSRC, "stone_width.cpp"),
                    os.path.join(CSRC, "host_scene.cpp"), "-lpthread"], check=True)
    out = _run(exe, sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "asc", "*.asc"))))
    assert out["parsed"] > 100 and out["rejected"] > 100 and out["solids"] > 100
    assert out["srb_width"] == pytest.approx(2.0, abs=1e-6)


def test_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "sanitize_oracle")
    subprocess.run(["gcc", "-std=gnu11", *SAN, "-ffp-contract=off", "-fno-builtin", "-march=x86-64-v3", "-o", exe,
                    os.path.join(ROOT, "tests", "native", "sanitize_oracle.c"), os.path.join(ROOT, "oracle", "ixo.c"),
                    "-lm", "-lpthread"], check=True)
    out = _run(exe, [])
    assert out["frames_sum"] > 0 and out["width"] == pytest.approx(2.0, abs=1e-6)
