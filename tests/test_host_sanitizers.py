"""The library's host code (the .ini parser, loader, generator, check_data
port and snapshots) and the oracle built with AddressSanitizer and
UndefinedBehaviorSanitizer (-fno-sanitize-recover: any report fails) and
run by tests/host_sanitize.cpp on the fixture, a generated tessellation and
hundreds of corrupted meshes with ids in and out of range."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "se-grain-growth-gpu_amd", "csrc")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1",
       "-ffp-contract=off", "-fPIC"]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_under_asan_and_ubsan(tmp_path):
    objs = []
    for src, cc, std in ((os.path.join(CSRC, "sevc_host.cpp"), "g++", "-std=c++17"),
                         (os.path.join(CSRC, "sevc_voronoi.cpp"), "g++", "-std=c++17"),
                         (os.path.join(REPO, "oracle", "sevc_oracle.c"), "gcc", "-std=c11"),
                         (os.path.join(REPO, "oracle", "sevc_oracle_io.c"), "gcc", "-std=c11"),
                         (os.path.join(REPO, "tests", "host_sanitize.cpp"), "g++", "-std=c++17")):
        o = str(tmp_path / (os.path.basename(src) + ".o"))
        subprocess.run([cc, std, "-Wno-unknown-pragmas", *SAN, "-c", src, "-o", o], check=True, capture_output=True)
        objs.append(o)
    exe = str(tmp_path / "host_sanitize")
    subprocess.run(["g++", *SAN, "-o", exe, *objs, "-lm"], check=True, capture_output=True)
    out = tmp_path / "out"
    out.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, os.path.join(REPO, "tests", "golden"), str(out)], capture_output=True, text=True,
                       env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "ALL OK" in r.stdout
