"""Builds and runs tests/host/host_unit.cpp: the C++ host layer's option parsing, format
IO, plain scheduler and slot placement (no GPU calls)."""
import os
import subprocess

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_unit(tmp_path):
    exe = str(tmp_path / "host_unit")
    build = os.path.join(REPO, "gowalker_amd", "_build")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(REPO, "include"), "-I",
                           os.path.join(REPO, "gowalker_amd", "host"), "-o", exe,
                           os.path.join(REPO, "tests", "host", "host_unit.cpp"), "-L", build, "-lgw",
                           "-Wl,-rpath," + build])
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=60)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "all passed" in r.stdout
