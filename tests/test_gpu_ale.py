"""Lagrangian / ALE mode (ale_positions) through the C ABI against the
oracle: moving positions, a bit-exact search every step, FP64 fields."""
import numpy as np
import pytest

import abi
import cases
import oracle
import urbansph
from test_gpu_parity import FIELDS, FP64_TOL, assert_fields, assert_nnps_equal, run_pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [abi.FP64])
def test_ale_box_fp64(precision):
    P, _ = cases.box(nf=10, walls=3, sigma_u=0.5)
    P["eli"] = np.where(P["p_type"] == 1, 1.0, 0.0).astype(np.float32)
    prm = cases.box_params(urbansph.default_params(), tol=1e-10)
    prm.update(ale_positions=1)
    o, g = run_pair(P, prm, steps=3, dt=2e-3)
    assert_nnps_equal(o, g)
    _, perm_o, _, _ = o.nnps()
    assert np.any(perm_o != np.arange(len(perm_o)))
    assert_fields(o, g, FP64_TOL, fields=FIELDS + ["X", "Y", "Z"])


def test_ale_mixed_runs():
    P, nrm, geom = cases.small_canyon(dx=0.02)
    P["eli"] = np.where(P["p_type"] == 1, 1.0, 0.0).astype(np.float32)
    prm = cases.canyon_params(geom, urbansph.default_params(), tol=1e-6, precision=abi.MIXED)
    prm.update(ale_positions=1)
    g = urbansph.Case(prm, len(P))
    g.upload(P, None, nrm)
    x0 = None
    for c in range(4):
        g.step(c, 5e-5)
        if c == 0:
            x0 = np.sort(g.field("X"))
    g.sync()
    r = g.ppe_report()
    assert r.converged == 1
    assert np.all(np.isfinite(g.field("PRES")))
    assert np.abs(np.sort(g.field("X")) - x0).max() > 0
