"""Batched stepping through the C ABI (sph_batch_*).

A case stepped inside a batch must give bitwise the same state as the same
case stepped alone with sph_step: segments never share a workgroup, each case
keeps its own Krylov scalars, tolerance and canonical reductions.  HIP-graph
replay of the Krylov iterations and the workgroup-size autotuner must not
change a single bit either.
"""
import numpy as np
import pytest

import abi
import cases
import urbansph
from test_gpu_parity import check_mixed

pytestmark = pytest.mark.gpu

FIELDS = ["UX", "UY", "UZ", "PRES", "TEMP", "VIS_T", "FLT_S", "FBX", "FBY", "FBZ", "FMX", "FMY", "HMZ",
          "CONCENTRATION", "SR", "RHO_REF", "UX0", "FTOTALX", "FTOTALZ", "FPX", "FPZ", "DTEMP", "INV_CM_XX",
          "INV_CM_XY"]


def _specs(precision):
    """Uneven cases with different physics: two boxes and two canyons (SSM,
    MOST, IBM buildings), each with its own dt."""
    out = []
    for k, nf in enumerate((7, 11)):
        P, nrm = cases.box(nf=nf, walls=3, seed_pos=10 + k, seed_vel=20 + k)
        prm = cases.box_params(urbansph.default_params(), tol=1e-8)
        prm.update(precision=precision)
        out.append((P, nrm, prm, 1e-4 * (1 + 0.25 * k)))
    for k, (P, nrm, geom, info) in enumerate(cases.ensemble(n_cases=2, lo=6e3, hi=2.5e4, seed=11, dx=0.01)):
        dt = 5e-5 * (1 + 0.2 * k)
        prm = cases.canyon_params(geom, urbansph.default_params(), dt=dt, tol=1e-6, precision=precision)
        out.append((P, nrm, prm, dt))
    return out


def _make(spec, graphs=True, autotune=True):
    P, nrm, prm, _ = spec
    c = urbansph.Case(prm, len(P))
    c.options(graphs=graphs, autotune=autotune)
    c.upload(P, None, nrm)
    return c


def _state(c):
    r = c.ppe_report()
    return {f: c.field(f) for f in FIELDS}, (r.iterations, r.refinements, r.converged, r.rel_residual)


def _alone(spec, steps, graphs=True, autotune=True):
    c = _make(spec, graphs, autotune)
    for s in range(steps):
        c.step(s, spec[3])
    c.sync()
    out = _state(c)
    c.close()
    return out


def _assert_same(a, b, what):
    fa, ra = a
    fb, rb = b
    assert ra == rb, f"{what}: PPE report {ra} != {rb}"
    for f in FIELDS:
        np.testing.assert_array_equal(fa[f], fb[f], err_msg=f"{what}: {f}")


@pytest.mark.parametrize("precision", [abi.FP64, abi.MIXED])
def test_batch_equals_alone(precision):
    specs = _specs(precision)
    steps = 3
    print(f"[batch] precision={precision} sizes={[len(s[0]) for s in specs]}", flush=True)
    alone = [_alone(s, steps) for s in specs]
    b = urbansph.Batch()
    cs = [_make(s) for s in specs]
    for st in range(steps):
        b.step(cs, [st] * len(cs), [s[3] for s in specs])
    b.sync()
    for k, c in enumerate(cs):
        _assert_same(_state(c), alone[k], f"case {k} (n={len(specs[k][0])})")
        c.close()
    b.close()


def _specs_mg():
    """The MIXED cases with the multigrid preconditioner asked for: the
    7^3 box (343 unknowns) runs Jacobi, the others multigrid."""
    out = _specs(abi.MIXED)
    for _, _, prm, _ in out:
        prm.update(ppe_precond=abi.PRECOND_MULTIGRID)
    return out


def test_batch_equals_alone_multigrid():
    """Multigrid and Jacobi cases in one batch: every case's V-cycles run
    in launches shared with the others, and each case ends bitwise where it
    ends alone."""
    specs = _specs_mg()
    steps = 3
    print(f"[batch-mg] sizes={[len(s[0]) for s in specs]}", flush=True)
    alone = [_alone(s, steps) for s in specs]
    b = urbansph.Batch()
    cs = [_make(s) for s in specs]
    for st in range(steps):
        b.step(cs, [st] * len(cs), [s[3] for s in specs])
    b.sync()
    pcs = [c.ppe_report().precond for c in cs]
    print(f"[batch-mg] preconditioners {pcs}", flush=True)
    assert pcs[0] == abi.PRECOND_JACOBI and all(p == abi.PRECOND_MULTIGRID for p in pcs[1:])
    for k, c in enumerate(cs):
        _assert_same(_state(c), alone[k], f"case {k} (n={len(specs[k][0])})")
        c.close()
    b.close()


def test_batch_staggered_membership_multigrid():
    """As test_batch_staggered_membership with the multigrid: the V-cycle
    tables follow the batch's membership."""
    specs = _specs_mg()
    steps = [4, 3, 2, 4]
    join = [0, 1, 2, 2]
    alone = [_alone(s, steps[k]) for k, s in enumerate(specs)]
    b = urbansph.Batch()
    cs = [_make(s) for s in specs]
    done = [0] * len(specs)
    for t in range(8):
        members = [k for k in range(len(specs)) if join[k] <= t and done[k] < steps[k]]
        if not members:
            break
        print(f"[batch-mg] tick {t}: members {members} counts {[done[k] for k in members]}", flush=True)
        b.step([cs[k] for k in members], [done[k] for k in members], [specs[k][3] for k in members])
        for k in members:
            done[k] += 1
    b.sync()
    for k, c in enumerate(cs):
        assert done[k] == steps[k]
        _assert_same(_state(c), alone[k], f"case {k}")
        c.close()
    b.close()


def test_batch_staggered_membership():
    """Cases join and leave the batch at different step counters (the
    serving pattern); each still matches its solo run bitwise."""
    specs = _specs(abi.MIXED)
    steps = [4, 3, 2, 4]
    join = [0, 1, 2, 2]
    alone = [_alone(s, steps[k]) for k, s in enumerate(specs)]
    b = urbansph.Batch()
    cs = [_make(s) for s in specs]
    done = [0] * len(specs)
    for t in range(8):
        members = [k for k in range(len(specs)) if join[k] <= t and done[k] < steps[k]]
        if not members:
            break
        print(f"[batch] tick {t}: members {members} counts {[done[k] for k in members]}", flush=True)
        b.step([cs[k] for k in members], [done[k] for k in members], [specs[k][3] for k in members])
        for k in members:
            done[k] += 1
    b.sync()
    for k, c in enumerate(cs):
        assert done[k] == steps[k]
        _assert_same(_state(c), alone[k], f"case {k}")
        c.close()
    b.close()


@pytest.mark.parametrize("precision", [abi.FP64, abi.MIXED])
def test_graphs_and_autotune_change_nothing(precision):
    """Steps 1-4 time the autotuner's candidates; from step 5 on the
    physics before and after the solve runs directly once, is captured at
    the next step and replayed after that: ten steps cover every stage."""
    spec = _specs(precision)[3]
    steps = 10
    ref = _alone(spec, steps, graphs=False, autotune=False)
    _assert_same(_alone(spec, steps, graphs=True, autotune=False), ref, "graphs")
    _assert_same(_alone(spec, steps, graphs=False, autotune=True), ref, "autotune")
    _assert_same(_alone(spec, steps, graphs=True, autotune=True), ref, "graphs+autotune")


def test_autotuner_picks_sizes():
    spec = _specs(abi.FP64)[2]
    c = _make(spec)
    for s in range(6):
        c.step(s, spec[3])
    c.sync()
    t = c.tuning()
    assert t, "no size class tuned"
    for cls, row in t.items():
        assert 2 ** cls <= len(spec[0]) < 2 ** (cls + 1)
        for name, blk in row.items():
            assert blk in (64, 128, 256, 512), (name, blk)
    c.close()


def test_batch_rejects_mixed_precisions():
    s64 = _specs(abi.FP64)[0]
    smx = _specs(abi.MIXED)[1]
    b = urbansph.Batch()
    a, c = _make(s64), _make(smx)
    with pytest.raises(urbansph.SPHError) as e:
        b.step([a, c], [0, 0], [1e-4, 1e-4])
    assert e.value.rc == abi.E_ARG
    with pytest.raises(urbansph.SPHError) as e:
        b.step([a, a], [0, 0], [1e-4, 1e-4])
    assert e.value.rc == abi.E_ARG
    a.close()
    c.close()
    b.close()


def _alone_input_order(spec, steps):
    c = _make(spec)
    for s in range(steps):
        c.step(s, spec[3])
    c.sync()
    perm = c.nnps()[1]
    out = {}
    for f in ("UX", "UY", "UZ", "PRES"):
        v = np.empty(c.n)
        v[perm] = c.field(f)
        out[f] = v
    c.close()
    return out


def test_server_results_equal_alone():
    """Requests arriving while others run are admitted as slots free up,
    co-batched, and each finishes with bitwise the result it has alone."""
    specs = _specs(abi.MIXED)
    steps = [3, 2, 4, 3]
    alone = [_alone_input_order(s, steps[k]) for k, s in enumerate(specs)]
    srv = urbansph.Server(budget_bytes=1 << 36, max_cases=3)
    ids = {}
    pending = list(range(len(specs)))
    done = []
    for t in range(40):
        if pending:   # one new request per tick, while others run
            k = pending.pop(0)
            P, nrm, prm, dt = specs[k]
            ids[k] = srv.submit(prm, P, steps[k], dt, normals=nrm)
        st = srv.status()
        print(f"[server] tick {t}: queued {st.queued} running {st.running} finished {st.finished}", flush=True)
        done += srv.tick()
        if not pending and srv.status().running == 0 and srv.status().queued == 0:
            break
    assert sorted(done) == sorted(ids.values())
    for k, rid in ids.items():
        i = srv.request(rid)
        assert i.state == abi.REQ_DONE and i.steps_done == steps[k] and i.rc == 0
        assert i.t_arrival <= i.t_admit <= i.t_finish
        for f in ("UX", "UY", "UZ", "PRES"):
            np.testing.assert_array_equal(srv.result(rid, f, len(specs[k][0])), alone[k][f], err_msg=f"case {k} {f}")
        p1 = srv.result_part1(rid, len(specs[k][0]))
        np.testing.assert_array_equal(p1["pres"], alone[k]["PRES"].astype(np.float32))
        np.testing.assert_array_equal(p1["x"], specs[k][0]["x"])
    st = srv.status()
    assert st.reserved_bytes == 0 and st.running == 0
    srv.close()


def test_index16_changes_nothing():
    """16-bit PPE column indices (MIXED) give bitwise the 32-bit result."""
    spec = _specs(abi.MIXED)[3]
    P, nrm, prm, dt = spec
    prm32 = type(prm).from_buffer_copy(prm)
    prm32.update(ppe_index16=0)
    ref = _alone((P, nrm, prm32, dt), 3)
    _assert_same(_alone(spec, 3), ref, "index16")


def test_lattice_order_matches_cell_order():
    """MIXED with the PPE unknowns numbered in lattice order is as close to a
    tightly converged FP64 solve as the cell-order numbering is (both solve
    to the same PPE tolerance; only their summation orders differ)."""
    spec = _specs(abi.MIXED)[3]
    P, nrm, prm, dt = spec
    prm0 = type(prm).from_buffer_copy(prm)
    prm0.update(ppe_reorder=0)
    ref_prm = type(prm).from_buffer_copy(prm)
    ref_prm.update(precision=abi.FP64, ppe_tol=1e-11)
    tol_prm = type(prm).from_buffer_copy(prm)
    tol_prm.update(precision=abi.FP64)
    ref, _ = _alone((P, nrm, ref_prm, dt), 3)
    at_tol, _ = _alone((P, nrm, tol_prm, dt), 3)
    a, ra = _alone((P, nrm, prm0, dt), 3)
    b, rb = _alone(spec, 3)
    assert ra[2] == 1 and rb[2] == 1
    for name, got, p in (("cell", a, prm0), ("lattice", b, prm)):
        tight = type(p).from_buffer_copy(p)
        tight.update(ppe_tol=1e-10)
        t, _ = _alone((P, nrm, tight, dt), 3)
        check_mixed(f"lattice {name}", got, ref, at_tol, t)


@pytest.mark.parametrize("precision,precond,tiles", [(abi.MIXED, abi.PRECOND_MULTIGRID, 0),
                                                     (abi.FP64, abi.PRECOND_MULTIGRID, 0),
                                                     (abi.MIXED, abi.PRECOND_JACOBI, 0), (abi.FP64, abi.PRECOND_JACOBI, 0),
                                                     (abi.MIXED, abi.PRECOND_MULTIGRID, 4 | 4 << 8 | 2 << 16)])
def test_admission_estimate_bounds_device_memory(precision, precond, tiles):
    """The server admits by sph_estimate_bytes before a request has run; after
    two steps the context holds no more than that (live and cached blocks,
    the multigrid set-up and the cell tiles included), and the estimate is
    not loose by more than 3x."""
    P, nrm, geom = cases.sized_canyon(1.2e5)
    prm = cases.canyon_params(geom, urbansph.default_params(), tol=1e-6, precision=precision)
    prm.update(ppe_precond=precond, tile_cells=tiles)
    est = urbansph.estimate_bytes(prm, len(P))
    c = urbansph.Case(prm, len(P))
    c.upload(P, None, nrm)
    for s in range(2):
        c.step(s, 5e-5)
    c.sync()
    held = c.device_bytes()
    pc = c.ppe_report().precond
    c.close()
    print(f"[estimate] precision {precision} precond {pc} tiles {tiles:#x}: held {held / 1e6:.1f} MB, "
          f"estimate {est / 1e6:.1f} MB",
          flush=True)
    assert pc == precond
    assert held <= est and est <= 3 * held
