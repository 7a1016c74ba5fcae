"""GPU farthest-point sampling (gp_farthest_point_sampling_ptp_gpu) against
the oracle's restatement of farthest_point_sampling_ptp_gpu
(src/cuda/geodesics_ptp.cu:87-172).  In FP64 every round's field is the
reference's, so the sample sequence and the returned radius are identical;
in FP32 the radius is within 1e-4 of the FP64 one."""
import numpy as np
import pytest

from gproshan_amd import _lib as L
from gproshan_amd import geodesics as P
from gproshan_amd import meshgen as G

pytestmark = pytest.mark.gpu

def _two_parts():
    """Two disjoint height fields: the first rounds reach only the part of the
    first sample, so the farthest vertex is the other part's first (its
    distance is +inf)."""
    x1, f1 = G.height_field(32, 5)
    x2, f2 = G.height_field(24, 6)
    x2 = x2 + np.array([3.0, 0.0, 0.0])
    return np.vstack([x1, x2]), np.vstack([f1, f2 + len(x1)])


CASES = [
    ("ico4", lambda: G.icosphere(4, 0.01, 2), [0], 24, 0.0),
    ("two_parts", _two_parts, [40], 12, 0.0),
    ("hf64", lambda: G.height_field(64, 9), [100], 16, 0.0),
    ("hf96_3init", lambda: G.height_field(96, 3), [5, 4000, 9000], 20, 0.0),
    ("hf64_radius", lambda: G.height_field(64, 11), [0], 200, 0.2),
    ("polar24", lambda: G.polar_disk(20, 24, 2), [0], 16, 0.0),
    ("polar10", lambda: G.polar_disk(20, 10, 3), [7], 16, 0.0),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fps_f64_matches_oracle(case, oracle):
    name, gen, init, n, radio = case
    xyz, f = gen()
    M = oracle.Mesh(xyz, f)
    ref, rref = M.fps(init, n, radio)
    C = P.che(xyz, f)
    got, r, secs = P.farthest_point_sampling_ptp_gpu(C, init, n, radio, precision=L.PREC_F64)
    print(f"\n[fps] {name} {len(got)} samples radio {r:.6f} (oracle {rref:.6f}) {secs * 1e3:.1f} ms", flush=True)
    assert got.tolist() == ref.tolist()
    assert r == rref
    g32, r32, _ = P.farthest_point_sampling_ptp_gpu(C, init, n, radio, precision=L.PREC_F32)
    print(f"[fps] {name} f32 radio {r32:.6f}, first difference at "
          f"{next((k for k in range(min(len(g32), len(got))) if g32[k] != got[k]), None)}", flush=True)
    assert len(g32) == len(got) or radio > 0
    if radio == 0:
        assert abs(r32 - r) <= 1e-4 * r


def test_fps_errors_and_edges():
    xyz, f = G.icosphere(2, 0.0, 1)
    C = P.che(xyz, f)
    with pytest.raises(P.GPError) as e:
        P.farthest_point_sampling_ptp_gpu(C, [3, 3], 10)
    assert e.value.status == L.ERR_DUPLICATE_SOURCE
    s, r, _ = P.farthest_point_sampling_ptp_gpu(C, [0, 1], 2)     # nothing to add
    assert s.tolist() == [0, 1] and r == np.inf
    s, r, _ = P.farthest_point_sampling_ptp_gpu(C, [0], 10 ** 9)  # V/2 in all
    assert len(s) == len(xyz) // 2 and len(set(s.tolist())) == len(s)
