"""Plans from several host threads (each plan owns its stream): the first GPU
use of a shared mesh (its upload and vertex-major tables) may happen in
several threads at once, and the fields equal those of sequential solves."""
import threading

import numpy as np
import pytest

from gproshan_amd import _lib as L
from gproshan_amd import geodesics as P
from gproshan_amd import meshgen as G

pytestmark = pytest.mark.gpu

SOURCES = [[100], [5000], [20000], [30000], [100, 30000], [777]]
KINDS = [dict(precision=L.PREC_F64), dict(precision=L.PREC_F64, device_setup=True),
         dict(precision=L.PREC_F32, device_setup=True), dict(precision=L.PREC_F32, uncapped=True, tol=1e-6)]


def test_plans_from_threads():
    xyz, f = G.height_field(192, 40)
    ref_mesh = P.che(xyz, f)
    ref = {}
    for q, s in enumerate(SOURCES):
        for k, kw in enumerate(KINDS):
            plan = P.Plan(ref_mesh, s, **kw)
            plan.run()
            ref[q, k] = plan.result()
            plan.close()
    mesh = P.che(xyz, f)   # fresh handle: its first GPU use happens in the threads
    out, errors = {}, []

    def work(q):
        try:
            for rep in range(2):
                for k, kw in enumerate(KINDS):
                    plan = P.Plan(mesh, SOURCES[q], **kw)
                    plan.run()
                    out[q, k, rep] = plan.result()
                    plan.close()
        except Exception as e:   # reported by the main thread
            errors.append((q, repr(e)))

    threads = [threading.Thread(target=work, args=(q,)) for q in range(len(SOURCES))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "a worker did not finish"
    assert not errors, errors
    for (q, k, rep), d in out.items():
        assert np.array_equal(d, ref[q, k], equal_nan=True), (q, k, rep)
    assert len(out) == len(SOURCES) * len(KINDS) * 2
