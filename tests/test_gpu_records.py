"""Compact MIXED records against the 128-byte double records and field arrays.

With float positions (uploaded, or refined children rounded to float), the
MIXED gathering passes read 32- and 64-byte float records of the neighbours,
and the Neumann pass 48-byte records of its pairs' fluid set;
compact_records = 0 makes them read the 128-byte double records and the
field arrays instead.  The float pair math sees the same values either way
((float)(x_i - x_j) of two float values is the rounded exact difference; the
other fields enter rounded to float), so every field must come out bitwise
the same: on the canyon with the multigrid over several steps, and on an
APR-refined canyon (heavy lists, long PPE rows).
"""
import numpy as np
import pytest

import abi
import cases
import urbansph
from test_gpu_apr import _refined
from test_gpu_tiles import _canyon, _run, _same

pytestmark = pytest.mark.gpu


def test_compact_records_equal_double_records():
    P, nrm, prm = _canyon()
    assert prm.compact_records == 1   # the default
    got, _ = _run(P, nrm, prm, steps=4)
    ref, _ = _run(P, nrm, type(prm).from_buffer_copy(prm).update(compact_records=0), steps=4)
    _same(got, ref, "compact records")
    print(f"[records] n={len(P)}: compact and double records agree bitwise over 4 steps", flush=True)


@pytest.mark.parametrize("precond", [abi.PRECOND_JACOBI, abi.PRECOND_MULTIGRID])
def test_compact_records_equal_double_records_refined(precond):
    _, _, geom = cases.small_canyon(dx=0.02)
    prm = cases.canyon_params(geom, urbansph.default_params(), tol=1e-6, precision=abi.MIXED)
    prm.update(ppe_precond=precond)
    got, live, r, n, n0 = _refined(prm)
    ref, live_r, r_r, n_r, _ = _refined(type(prm).from_buffer_copy(prm).update(compact_records=0))
    assert n == n_r and n > n0 and np.array_equal(live, live_r)
    assert (r.iterations, r.refinements, r.rel_residual) == (r_r.iterations, r_r.refinements, r_r.rel_residual)
    for f in got:
        np.testing.assert_array_equal(got[f], ref[f], err_msg=f)
    print(f"[records] refined n {n0} -> {n}, precond {precond}: compact and double records agree bitwise",
          flush=True)
