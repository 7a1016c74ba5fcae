"""Builds and runs the C++ drop-in front-end example (tests/cpp/frontend.cpp)
against libgproshan_ptp.so: PTP_CPU on the CPU, PTP_GPU with the gpu marker."""
import os
import subprocess

import numpy as np
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


def _grid20():
    """The 20 x 20 planar grid of tests/cpp/frontend.cpp (same vertex order
    and CCW triangles)."""
    n = 20
    xyz = np.array([(c / (n - 1.0), r / (n - 1.0), 0.0) for r in range(n) for c in range(n)])
    f = []
    for r in range(n - 1):
        for c in range(n - 1):
            a = r * n + c
            b, d = a + 1, a + n
            e = d + 1
            f += [a, b, e, a, e, d]
    return xyz, np.array(f, np.uint32), (n // 2) * n + n // 2


@pytest.mark.gpu
def test_frontend_gpu(exe, oracle, tmp_path):
    prefix = str(tmp_path / "ptp_gpu")
    r = subprocess.run([exe, "gpu", prefix], capture_output=True, text=True)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "OK" in r.stdout
    # PTP_GPU of the C++ front end against the FP64 oracle of the reference's
    # one-source GPU path (src/geodesics.cpp:232-235: the coalescence driver,
    # last-written buffer): FP64 bit-exact, FP32 within 1e-4.
    xyz, f, s = _grid20()
    ref, _ = oracle.Mesh(xyz, f).ptp([s], oracle.GPU_COAL)
    d64 = np.fromfile(prefix + ".f64", np.float64)
    d32 = np.fromfile(prefix + ".f32", np.float64)
    assert np.array_equal(d64, ref)
    m = ref > 0
    assert np.max(np.abs(d32[m] - ref[m]) / ref[m]) <= 1e-4
