"""GPU iteration-error instrumentation against the oracle: FP64 band rows and
sweep numbers identical, errors equal to summation order (1e-10)."""
import numpy as np
import pytest

from gproshan_amd import _lib as L
from gproshan_amd import geodesics as P
from gproshan_amd import meshgen as G

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,seed,src", [(96, 6, [200]), (128, 1, [777, 5000, 9000])])
def test_iter_error_f64_matches_oracle(oracle, n, seed, src):
    xyz, f = G.height_field(n, seed)
    M = oracle.Mesh(xyz, f)
    src = np.array(src, np.uint32)
    exact = M.fixed_point(src)
    it0, err0, band0 = M.iter_error(src, exact)
    C = P.che(xyz, f)
    it, err, band = P.iter_error_parallel_toplesets_propagation_coalescence_gpu(C, src, exact, L.PREC_F64)
    print(f"\n[iter_error] {n}^2 {len(src)} sources: {len(band)} sweeps, {len(it)} errors, last {err[-1]:.4e} %",
          flush=True)
    assert np.array_equal(band, band0)
    assert np.array_equal(it, it0)
    assert np.allclose(err, err0, rtol=1e-10, atol=0)
    # err is a mean relative error in percent: an FP32 field whose mean
    # relative deviation from the FP64 one is 1e-7 (about two FP32 ulps of
    # accumulated rounding) moves it by at most 1e-5.
    it32, err32, _ = P.iter_error_parallel_toplesets_propagation_coalescence_gpu(C, src, exact, L.PREC_F32)
    print(f"  f32 last {err32[-1]:.6e} % (f64 {err0[-1]:.6e} %)", flush=True)
    assert abs(err32[-1] - err0[-1]) <= 1e-5
