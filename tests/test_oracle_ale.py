"""Lagrangian / ALE mode of the oracle (ale_positions): the corrector moves a
fluid particle by the trapezoid of its start and corrected velocities times
eli (functions_TIME_ISPH.cuh:193-195), walls stay, and the next step's
search bins the moved positions."""
import numpy as np

import abi
import cases
import oracle

f32 = np.float32


def _ale_box(ale=1, eli=1.0):
    P, _ = cases.box(nf=8, walls=3, sigma_u=0.5)
    P["eli"] = np.where(P["p_type"] == 1, eli, 0.0).astype(np.float32)
    prm = cases.box_params(oracle.default_params())
    prm.update(ale_positions=ale)
    o = oracle.Oracle(prm, len(P))
    o.upload(P)
    return P, o


def test_corrector_moves_fluid_by_the_velocity_trapezoid():
    P, o = _ale_box()
    dt = 1e-4
    assert o.step(0, dt) == 0
    _, perm, _, _ = o.nnps()
    x0 = P["x"].astype(np.float64)[perm]
    y0 = P["y"].astype(np.float64)[perm]
    pt = o.ifield("P_TYPE")
    mv = (pt > 0).astype(np.float64) * P["eli"].astype(np.float64)[perm]
    ux0, ux = o.field("UX0"), o.field("UX")
    uy0, uy = o.field("UY0"), o.field("UY")
    np.testing.assert_array_equal(o.field("X"), x0 + (ux0 + ux) / 2.0 * dt * mv)
    np.testing.assert_array_equal(o.field("Y"), y0 + (uy0 + uy) / 2.0 * dt * mv)
    walls = pt == 0
    np.testing.assert_array_equal(o.field("X")[walls], x0[walls])
    assert np.abs(o.field("X") - x0)[~walls].max() > 0


def test_positions_fixed_without_ale_or_eli():
    for ale, eli in ((0, 1.0), (1, 0.0)):
        P, o = _ale_box(ale, eli)
        assert o.step(0, 1e-4) == 0
        _, perm, _, _ = o.nnps()
        np.testing.assert_array_equal(o.field("X"), P["x"].astype(np.float64)[perm])


def test_next_step_searches_the_moved_positions():
    P, o = _ale_box()
    dt = 2e-3        # large enough that particles change cells
    assert o.step(0, dt) == 0
    moved = {k: o.field(k) for k in ("X", "Y", "Z")}
    it = o.ifield("I_TYPE")
    assert o.step(1, dt) == 0
    d = o.domain()
    keys, perm, _, _ = o.nnps()

    def axis(v, mn, mx, N):
        t = np.floor((v.astype(f32) - f32(mn)) / (f32(mx) - f32(mn)) * f32(N))
        return np.maximum(np.minimum(t, f32(N - 1)).astype(np.int64), 0)

    c = np.stack([axis(moved["X"], d.x_min, d.x_max, d.NI), axis(moved["Y"], d.y_min, d.y_max, d.NJ),
                  axis(moved["Z"], d.z_min, d.z_max, d.NK)], 1)
    key = c[:, 0] + d.NI * c[:, 1] + d.NI * d.NJ * c[:, 2]
    key[it == 3] = d.num_cells
    exp = np.argsort(key, kind="stable")
    np.testing.assert_array_equal(perm, exp)
    np.testing.assert_array_equal(keys, key[exp])
    assert np.any(exp != np.arange(len(exp)))   # the search reordered the particles
