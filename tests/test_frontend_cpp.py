"""Builds and runs the C++ drop-in front-end example (tests/cpp/frontend.cpp)
against libgproshan_ptp.so: PTP_CPU on the CPU, PTP_GPU with the gpu marker."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "frontend")


@pytest.fixture(scope="module")
def exe():
    src = os.path.join(ROOT, "tests", "cpp", "frontend.cpp")
    lib = os.path.join(ROOT, "gproshan_amd")
    deps = [src, os.path.join(ROOT, "include", "gproshan", "geodesics.h"), os.path.join(ROOT, "include", "gproshan_ptp.h")]
    if os.path.exists(EXE) and all(os.path.getmtime(EXE) >= os.path.getmtime(d) for d in deps):
        return EXE
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ROOT}/include", src, "-o", EXE, f"-L{lib}", "-lgproshan_ptp",
                    f"-Wl,-rpath,{lib}"], check=True)
    return EXE


def test_frontend_cpu(exe):
    r = subprocess.run([exe], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0 and "OK" in r.stdout


@pytest.mark.gpu
def test_frontend_gpu(exe):
    r = subprocess.run([exe, "gpu"], capture_output=True, text=True)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "OK" in r.stdout
