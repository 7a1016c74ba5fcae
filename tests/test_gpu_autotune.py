"""The launch-shape and band-width autotuners (gp_plan_autotune,
gp_plan_create_bands_tuned).

  * tuning changes the launch shape only: FP64 fields and schedule traces are
    bit-identical before and after, the chosen grid and replay length are among
    the candidates, and a second plan of the same shape takes the cached choice
    without timing;
  * the tuned band plan's width is one of the candidates and its field is
    within 1e-4 of the FP64 fixed point of the relaxation.
"""
import numpy as np
import pytest

from conftest import rel_err
from gproshan_amd import _lib as L
from gproshan_amd import geodesics as P
from gproshan_amd import meshgen as G

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kw", [dict(), dict(uncapped=True, tol=1e-6)], ids=["ref", "uncapped"])
def test_autotune_keeps_results(kw):
    xyz, f = G.height_field(200, 21)
    C = P.che(xyz, f)
    src = [12345]
    plan = P.Plan(C, src, precision=L.PREC_F64, **kw)
    plan.run()
    d0, tr0 = plan.result(), plan.trace()
    t = plan.autotune()
    print(f"\n[autotune] {kw}: {t}", flush=True)
    assert t["candidates"] >= 4 and t["replay"] in (16, 32, 64, 128) and t["grid"] >= 8 and t["kernel_ms"] > 0
    st = plan.run()
    assert np.array_equal(plan.result().view(np.uint64), d0.view(np.uint64))
    assert np.array_equal(plan.trace(), tr0)
    plan.close()
    again = P.Plan(C, [777], precision=L.PREC_F64, **kw)
    t2 = again.autotune()
    assert t2["candidates"] == 0 and t2["grid"] == t["grid"] and t2["replay"] == t["replay"]
    again.close()


def test_band_width_tuner(oracle):
    xyz, f = G.height_field(192, 22)
    C = P.che(xyz, f)
    src = [9000]
    fp = oracle.Mesh(xyz, f).fixed_point(src)
    e = np.linalg.norm(xyz - xyz[src[0]], axis=1)
    widths = [0.5, 0.75, 1.0]
    plan, w = P.Plan.tuned_bands(C, src, e, widths=widths, schedule=L.SCHED_EUCLID_BAND, tol=1e-6)
    st = plan.run()
    mx, _ = rel_err(plan.result(), fp)
    print(f"\n[band tuner] width {w} sweeps {st['sweeps']} + {st['check_sweeps']} check, max {mx:.2e}", flush=True)
    plan.close()
    assert w in widths
    assert mx <= 1e-4


_CHILD = r"""
import json, sys
sys.path.insert(0, sys.argv[1])
from gproshan_amd import _lib as L, geodesics as P, meshgen as G
xyz, f = G.height_field(96, 23)
plan = P.Plan(P.che(xyz, f), [4000], precision=L.PREC_F32, uncapped=True, tol=1e-6)
t = plan.autotune()
plan.close()
print(json.dumps(t))
"""


def _tune_in_child(tmp_path, tune_file):
    import json
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, GPROSHAN_AMD_TUNE_FILE=str(tune_file))
    r = subprocess.run([sys.executable, "-c", _CHILD, root], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_tune_file_persists_across_processes(tmp_path):
    """GPROSHAN_AMD_TUNE_FILE: the first process times the candidates and
    appends its choice; a second process reads it and times nothing; lines a
    damaged file holds (unknown kind, out-of-range shape) are not taken."""
    tf = tmp_path / "tune.txt"
    first = _tune_in_child(tmp_path, tf)
    print(f"\n[tune file] first process: {first}; file: {tf.read_text().strip()}", flush=True)
    assert first["candidates"] > 0
    second = _tune_in_child(tmp_path, tf)
    assert second["candidates"] == 0
    assert (second["grid"], second["replay"], second["xcd_block"]) == (first["grid"], first["replay"], first["xcd_block"])
    # A file whose only entry for this shape is out of range: tuned again.
    key = tf.read_text().split()[1:8]
    bad = tmp_path / "bad.txt"
    bad.write_text("shape " + " ".join(key) + " 4096 1000000 0 1.0\n")
    third = _tune_in_child(tmp_path, bad)
    assert third["candidates"] > 0 and third["replay"] in (16, 32, 64, 128)
