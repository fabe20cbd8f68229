"""sph_stats_noalloc (working memory in a library-owned device array, no
allocation) against sph_stats (the context's buffer), through the C ABI:
the same three maxima bit for bit, on two calls in a row, the second on a
smaller case whose maxima are below the first's, so a fold that read the
first call's leftover partials would return the first call's values."""
import numpy as np
import pytest

import cases
import urbansph

pytestmark = pytest.mark.gpu


def _stepped(nf, sigma_u):
    P, _ = cases.box(nf=nf, walls=3, sigma_u=sigma_u)
    prm = cases.box_params(urbansph.default_params(), tol=1e-8)
    g = urbansph.Case(prm, len(P))
    g.upload(P)
    g.step(0, 1e-4)
    g.sync()
    return g


def test_stats_noalloc_equals_stats_on_two_calls():
    big = _stepped(16, 0.1)       # 10,648 particles, |u| ~ 0.1 m/s
    small = _stepped(8, 0.01)     # 2,744 particles, |u| ~ 0.01 m/s
    for g, what in ((big, "first call, 10648 particles"), (small, "second call, 2744 particles")):
        print(f"[no_alloc] {what}", flush=True)
        before = g.device_bytes()
        got = np.array(g.stats_noalloc(), np.float32)
        assert g.device_bytes() == before, "the allocation-free call changed the context's device memory"
        ref = np.array(g.stats(), np.float32)
        assert np.isfinite(ref).all() and ref[0] > 0.0
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (what, got, ref)
    ub, us = big.stats()[0], small.stats()[0]
    assert ub > 2.0 * us, "the second case must have smaller maxima than the first"
    # the first case again, after the smaller one used the array
    assert np.array_equal(np.array(big.stats_noalloc(), np.float32).view(np.uint32),
                          np.array(big.stats(), np.float32).view(np.uint32))
