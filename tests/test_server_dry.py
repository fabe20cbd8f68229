"""Admission and continuous-batching policy of the serving loop, on the CPU.

The server's dry-run mode makes no GPU call: each tick advances a simulated
clock by (particles of the running cases) / rate.  These tests check the
policy itself: FIFO admission under the HBM budget and the slot limit,
requests joining the batch as soon as memory frees up, latency bookkeeping.
"""
import numpy as np
import pytest

import abi
import urbansph

RATE = 1.0e6   # modelled particle-steps per second


def _params(precision=abi.FP64):
    p = urbansph.default_params()
    p.update(ppe_solver=abi.PPE_BICGSTAB, precision=precision)
    return p


def _deck(n):
    return np.zeros(n, abi.PART1)


def test_estimate_scales_with_particles():
    p = _params()
    a, b = urbansph.estimate_bytes(p, 10_000), urbansph.estimate_bytes(p, 20_000)
    assert 0 < a < b
    # the per-particle part doubles, the fixed part does not
    assert abs((b - a) - (a - (16 << 20))) <= 1
    m = _params(abi.MIXED)
    assert urbansph.estimate_bytes(m, 10_000) < a   # float matrix and Krylov vectors
    # the multigrid preconditioner adds its hierarchy and set-up terms
    mg = _params(abi.MIXED)
    mg.update(ppe_precond=abi.PRECOND_MULTIGRID)
    assert urbansph.estimate_bytes(mg, 10_000) > urbansph.estimate_bytes(m, 10_000) + 10_000 * 1000
    # the cell tiles add their copy of the lists (MIXED only)
    t = _params(abi.MIXED)
    t.update(tile_cells=4 | 4 << 8 | 2 << 16)
    assert urbansph.estimate_bytes(t, 10_000) >= urbansph.estimate_bytes(m, 10_000) + 10_000 * 600
    f = _params()
    f.update(tile_cells=4 | 4 << 8 | 2 << 16)
    assert urbansph.estimate_bytes(f, 10_000) == a


def test_budget_admission_fifo_and_continuous_batching():
    p = _params()
    n = 10_000
    est = urbansph.estimate_bytes(p, n)
    srv = urbansph.Server(budget_bytes=2 * est + est // 2, max_cases=8, dry_run_rate=RATE)
    steps = [3, 5, 2, 4]
    ids = [srv.submit(p, _deck(n), s, 1e-4, t_arrival=0.0) for s in steps]
    finished_at = {}
    tick = 0
    while True:
        st = srv.status()
        if st.queued == 0 and st.running == 0:
            break
        done = srv.tick()
        tick += 1
        st = srv.status()
        assert st.reserved_bytes <= st.budget_bytes
        assert st.running <= 2                  # only two decks fit the budget
        for rid in done:
            finished_at[rid] = tick
        assert tick < 50
    info = {rid: srv.request(rid) for rid in ids}
    assert all(i.state == abi.REQ_DONE and i.steps_done == i.steps for i in info.values())
    # arrival order is admission order
    admits = [info[r].t_admit for r in ids]
    assert admits == sorted(admits)
    # request 1 (3 steps) and 2 (5 steps) start at once; 3 joins the tick
    # after 1 finishes, 4 the tick after 3 finishes
    assert finished_at[ids[0]] == 3
    assert finished_at[ids[2]] == 5
    assert info[ids[2]].t_admit == pytest.approx(info[ids[0]].t_finish)
    assert finished_at[ids[1]] == 5
    assert finished_at[ids[3]] == 9
    # latency = simulated time of the ticks it waited and ran through
    t_tick = 2 * n / RATE
    assert info[ids[0]].t_finish - info[ids[0]].t_arrival == pytest.approx(3 * t_tick)
    # two decks per tick through tick 5, then request 4 alone for four ticks
    assert info[ids[3]].t_finish == pytest.approx(5 * t_tick + 4 * n / RATE)
    st = srv.status()
    assert st.reserved_bytes == 0 and st.ticks == 9 and st.particle_steps == n * sum(steps)
    srv.close()


def test_slot_limit_and_idle_clock():
    p = _params()
    srv = urbansph.Server(budget_bytes=1 << 40, max_cases=3, dry_run_rate=RATE)
    ids = [srv.submit(p, _deck(1000), 2, 1e-4, t_arrival=0.0) for _ in range(5)]
    srv.tick()
    assert srv.status().running == 3 and srv.status().queued == 2
    while srv.status().running or srv.status().queued:
        srv.tick()
    # an idle server jumps to the next arrival
    srv.idle_until(10.0)
    late = srv.submit(p, _deck(1000), 1, 1e-4, t_arrival=10.0)
    srv.tick()
    i = srv.request(late)
    assert i.state == abi.REQ_DONE and i.t_admit == pytest.approx(10.0)
    assert i.t_finish - i.t_arrival == pytest.approx(1000 / RATE)
    for rid in ids + [late]:
        srv.release(rid)
    with pytest.raises(urbansph.SPHError):
        srv.request(late)
    srv.close()


def test_oversized_request_fails_without_blocking():
    p = _params()
    small = urbansph.estimate_bytes(p, 1000)
    srv = urbansph.Server(budget_bytes=3 * small, max_cases=8, dry_run_rate=RATE)
    big = srv.submit(p, _deck(100_000), 1, 1e-4)
    ok = srv.submit(p, _deck(1000), 1, 1e-4)
    assert srv.request(big).state == abi.REQ_FAILED and srv.request(big).rc == abi.E_NOMEM
    assert srv.tick() == [ok]
    srv.close()


def test_requests_share_precision():
    srv = urbansph.Server(budget_bytes=1 << 40, max_cases=4, dry_run_rate=RATE)
    srv.submit(_params(abi.FP64), _deck(100), 1, 1e-4)
    with pytest.raises(urbansph.SPHError) as e:
        srv.submit(_params(abi.MIXED), _deck(100), 1, 1e-4)
    assert e.value.rc == abi.E_ARG
    srv.close()


def test_poisson_load_latency_percentiles():
    """A seeded Poisson stream through the dry-run server: every request
    completes, latencies are at least the request's own run time, and p99 >=
    p50 (the bookkeeping the benchmark reports)."""
    import cases
    p = _params()
    rng_u = cases.uniform(5, 60).reshape(30, 2)
    t = 0.0
    arrivals = []
    for k in range(30):
        t += -np.log(1.0 - rng_u[k, 0]) / 4.0            # 4 requests / s
        n = int(np.exp(np.log(1e4) + rng_u[k, 1] * (np.log(2e5) - np.log(1e4))))
        arrivals.append((t, n, 5 + k % 7))
    srv = urbansph.Server(budget_bytes=int(2e9), max_cases=64, dry_run_rate=2e7)
    pending = list(arrivals)
    ids = []
    while pending or srv.status().running or srv.status().queued:
        while pending and pending[0][0] <= srv.clock():
            ta, n, s = pending.pop(0)
            ids.append((srv.submit(p, _deck(n), s, 1e-4, t_arrival=ta), n, s))
        if not srv.status().running and not srv.status().queued and pending:
            srv.idle_until(pending[0][0])
            continue
        srv.tick()
    lat = []
    for rid, n, s in ids:
        i = srv.request(rid)
        assert i.state == abi.REQ_DONE
        assert i.t_finish - i.t_arrival >= s * n / 2e7 * (1 - 1e-12)
        lat.append(i.t_finish - i.t_arrival)
    p50, p99 = np.percentile(lat, 50), np.percentile(lat, 99)
    assert 0 < p50 <= p99
    srv.close()
