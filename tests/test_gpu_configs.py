"""The benchmark configurations at their stated shapes, through the C ABI.

* the 32^3 FP64 box (54,872 particles, PPE tolerance 1e-8) against the
  oracle: bit-exact search, fields within 1e-10, the same iteration count;
* the 64-case canyon ensemble (seed 7, 2e4..3e5 particles each, MIXED with
  the multigrid): every case stepped in one batch ends bitwise where it ends
  alone, and its two smallest cases match the oracle;
* the serving mix (Poisson arrivals at 4/s, 1e4..2e6 particles, 50..500
  steps, L in {-50, +50, inf}): four requests, one of more than 1e6
  particles, each finishing with bitwise its result alone;
* a multigrid hierarchy of four levels against the converged FP64 oracle.
"""
import numpy as np
import pytest

import abi
import cases
import oracle
import urbansph
from test_gpu_batch import FIELDS, _alone, _alone_input_order, _assert_same, _make, _state
from test_gpu_parity import FP64_TOL, assert_fields, assert_nnps_equal, check_mixed, run_pair

pytestmark = pytest.mark.gpu

UP = ("PRES", "UX", "UY", "UZ")


def test_box_full_size_fp64():
    P, _ = cases.box(nf=32, walls=3)
    assert len(P) == 54872
    prm = cases.box_params(urbansph.default_params(), tol=1e-8)
    prm.update(ppe_solver=abi.PPE_BICGSTAB)
    o, g = run_pair(P, prm, steps=1)
    assert_nnps_equal(o, g)
    w = assert_fields(o, g, FP64_TOL, report={})
    ro, rg = o.ppe_report(), g.ppe_report()
    print(f"[cfg-box] iterations gpu {rg.iterations} oracle {ro.iterations}; worst rel L2 "
          f"{max(w.values()):.2e}", flush=True)
    assert rg.iterations == ro.iterations and rg.converged == 1


def _ensemble_specs():
    out = []
    for P, nrm, geom, info in cases.ensemble(n_cases=64, lo=2e4, hi=3e5, seed=7):
        prm = cases.canyon_params(geom, urbansph.default_params(), tol=1e-6, precision=abi.MIXED)
        prm.update(ppe_precond=abi.PRECOND_MULTIGRID)
        out.append((P, nrm, prm, 5e-5))
    return out


def test_ensemble_64_batch_equals_alone():
    specs = _ensemble_specs()
    sizes = [len(s[0]) for s in specs]
    print(f"[cfg-ensemble] 64 cases, {sum(sizes)} particles, {min(sizes)}..{max(sizes)}", flush=True)
    steps = 3
    b = urbansph.Batch()
    cs = [_make(s) for s in specs]
    for st in range(steps):
        b.step(cs, [st] * len(cs), [s[3] for s in specs])
    b.sync()
    got = [_state(c) for c in cs]
    pcs = [c.ppe_report().precond for c in cs]
    for c in cs:
        c.close()
    b.close()
    print(f"[cfg-ensemble] batch done; multigrid cases {sum(p == abi.PRECOND_MULTIGRID for p in pcs)}", flush=True)
    assert all(p == abi.PRECOND_MULTIGRID for p in pcs)
    for k, s in enumerate(specs):
        _assert_same(got[k], _alone(s, steps), f"case {k} (n={sizes[k]})")
    print("[cfg-ensemble] every case equals its solo run", flush=True)


def test_ensemble_smallest_cases_against_oracle():
    """The two smallest ensemble cases: FP64 against the oracle at the
    same tolerance (1e-10), and MIXED with the multigrid against the
    converged oracle (no farther from it than FP64 at the same tolerance)."""
    specs = _ensemble_specs()
    order = np.argsort([len(s[0]) for s in specs])[:2]
    for k in order:
        P, nrm, prm_mixed, dt = specs[k]
        prm64 = type(prm_mixed).from_buffer_copy(prm_mixed)
        prm64.update(precision=abi.FP64, ppe_precond=abi.PRECOND_JACOBI)
        o, g = run_pair(P, prm64, normals=nrm, steps=1, dt=dt)
        assert_nnps_equal(o, g)
        assert_fields(o, g, FP64_TOL)
        conv = type(prm64).from_buffer_copy(prm64)
        conv.update(ppe_tol=1e-12)
        oc = oracle.Oracle(conv, len(P))
        oc.upload(P, None, nrm)
        assert oc.step(0, dt) == 0
        gm = urbansph.Case(prm_mixed, len(P))
        gm.upload(P, None, nrm)
        gm.step(0, dt)
        gm.sync()
        assert gm.ppe_report().precond == abi.PRECOND_MULTIGRID
        got = {f: gm.field(f) for f in UP}
        gm.close()
        tight = type(prm_mixed).from_buffer_copy(prm_mixed)
        tight.update(ppe_tol=1e-10)
        gt = urbansph.Case(tight, len(P))
        gt.upload(P, None, nrm)
        gt.step(0, dt)
        gt.sync()
        check_mixed(f"cfg-ensemble case {k} n={len(P)}", got, {f: oc.field(f) for f in UP},
                    {f: o.field(f) for f in UP}, {f: gt.field(f) for f in UP})
        gt.close()


def test_serving_mix_full_size_equals_alone():
    reqs = cases.serving_requests(n_req=4, rate=4.0, lo=1e4, hi=2e6, steps_lo=50, steps_hi=500, seed=52)
    specs = []
    for r in reqs:
        P, nrm, geom = cases.sized_canyon(r["target"], r["u_scale"], r["bh"])
        prm = cases.canyon_params(geom, urbansph.default_params(), tol=1e-6, precision=abi.MIXED, L=r["L"])
        prm.update(ppe_precond=abi.PRECOND_MULTIGRID)
        specs.append((P, nrm, prm, 5e-5))
    steps = [r["steps"] for r in reqs]
    sizes = [len(s[0]) for s in specs]
    print(f"[cfg-serving] sizes {sizes} steps {steps} L {[r['L'] for r in reqs]}", flush=True)
    assert max(sizes) >= 1_000_000
    srv = urbansph.Server(budget_bytes=240 << 30, max_cases=256)
    # arrivals at the Poisson times, on a 20 ms tick clock: later requests
    # join while earlier ones are mid-run
    arrive = [int(round(r["t_arrival"] / 0.02)) for r in reqs]
    ids = {}
    t = 0
    while len(ids) < len(reqs) or srv.status().running or srv.status().queued:
        for k in range(len(reqs)):
            if k not in ids and arrive[k] <= t:
                P, nrm, prm, dt = specs[k]
                ids[k] = srv.submit(prm, P, steps[k], dt, normals=nrm)
        if srv.status().running or srv.status().queued:
            srv.tick()
        t += 1
        assert t < 5000
    overlap = srv.status().ticks
    print(f"[cfg-serving] {overlap} ticks for {sum(steps)} request-steps", flush=True)
    assert overlap < sum(steps)            # requests shared ticks
    for k, rid in ids.items():
        i = srv.request(rid)
        assert i.state == abi.REQ_DONE and i.steps_done == steps[k] and i.rc == 0
        alone = _alone_input_order(specs[k], steps[k])
        for f in UP:
            np.testing.assert_array_equal(srv.result(rid, f, sizes[k]), alone[f], err_msg=f"request {k} {f}")
        print(f"[cfg-serving] request {k} (n={sizes[k]}, {steps[k]} steps) equals its solo run", flush=True)
    srv.close()


def test_multigrid_four_levels_against_oracle():
    P, nrm, geom = cases.sized_canyon(2.6e5)
    n = len(P)
    dt = 5e-5
    conv = cases.canyon_params(geom, urbansph.default_params(), tol=1e-12)
    at6 = cases.canyon_params(geom, urbansph.default_params(), tol=1e-6)
    ref = {}
    for name, prm in (("conv", conv), ("fp64@1e-6", at6)):
        o = oracle.Oracle(prm, n)
        o.upload(P, None, nrm)
        assert o.step(0, dt) == 0
        ref[name] = {f: o.field(f) for f in UP}
    print(f"[cfg-mg4] n={n}: oracle done", flush=True)
    got = {}
    for tol in (1e-6, 1e-10):
        prm = cases.canyon_params(geom, urbansph.default_params(), tol=tol, precision=abi.MIXED)
        prm.update(ppe_precond=abi.PRECOND_MULTIGRID)
        g = urbansph.Case(prm, n)
        g.upload(P, None, nrm)
        g.step(0, dt)
        g.sync()
        r = g.ppe_report()
        print(f"[cfg-mg4] tol {tol:.0e}: levels {r.mg_levels} coarsest {r.mg_coarsest} iterations {r.iterations}",
              flush=True)
        assert r.precond == abi.PRECOND_MULTIGRID and r.mg_levels >= 4 and r.converged == 1
        got[tol] = {f: g.field(f) for f in UP}
        g.close()
    check_mixed("cfg-mg4", got[1e-6], ref["conv"], ref["fp64@1e-6"], got[1e-10])
