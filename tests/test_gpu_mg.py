"""Multigrid preconditioner of the MIXED inner solve (ppe_precond =
SPH_PRECOND_MULTIGRID), through the C ABI.

The preconditioner changes the Krylov path, not the system or its stopping
rule: at PPE tolerance 1e-6 the result must be as close to the converged
FP64 oracle as the FP64 solve at the same tolerance is, driven to a tight
tolerance it must stay within 1e-5 of it, it must take fewer iterations
than the Jacobi form, and it must be bit-reproducible (hierarchy build,
graphs or direct launches).  A PPE of at most 1024 unknowns runs the
Jacobi form, and the report says so; batches mixing both forms are covered
in test_gpu_batch.py.
"""
import numpy as np
import pytest

import abi
import cases
import oracle
import urbansph
from test_gpu_parity import check_mixed, mixed_errors

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    d = np.linalg.norm(a - b)
    s = np.linalg.norm(b)
    return d / s if s > 0 else d


def _case(which):
    if which == "box":
        P, nrm = cases.box(nf=12, walls=3)
        return P, nrm, lambda tol: cases.box_params(urbansph.default_params(), tol=tol), 1e-4
    P, nrm, geom = cases.small_canyon(dx=0.01)
    return P, nrm, lambda tol: cases.canyon_params(geom, urbansph.default_params(), tol=tol), 5e-5


def _oracle(P, nrm, prm, dt):
    o = oracle.Oracle(prm, len(P))
    o.upload(P, None, nrm)
    o.step(0, dt)
    out = {f: o.field(f) for f in ("PRES", "UX", "UY", "UZ")}
    o.close()
    return out


def _gpu(P, nrm, prm, dt, steps=1, precond=abi.PRECOND_MULTIGRID, graphs=True):
    prm.update(precision=abi.MIXED, ppe_precond=precond)
    print(f"[mg] n={len(P)} precond={precond} steps={steps} graphs={graphs}", flush=True)
    g = urbansph.Case(prm, len(P))
    g.options(graphs=graphs, autotune=True)
    g.upload(P, None, nrm)
    reps = []
    for c in range(steps):
        g.step(c, dt)
        g.sync()
        r = g.ppe_report()
        reps.append(dict(iterations=r.iterations, refinements=r.refinements, converged=r.converged,
                         rel_residual=r.rel_residual, precond=r.precond, mg_levels=r.mg_levels,
                         mg_coarsest=r.mg_coarsest, unknowns=r.unknowns))
    return g, reps


@pytest.mark.parametrize("which", ["box", "canyon"])
def test_multigrid_same_tolerance_as_fp64(which):
    P, nrm, mk, dt = _case(which)
    conv = _oracle(P, nrm, mk(1e-12), dt)
    at_tol = _oracle(P, nrm, mk(1e-6), dt)
    g, reps = _gpu(P, nrm, mk(1e-6), dt)
    _, jac = _gpu(P, nrm, mk(1e-6), dt, precond=abi.PRECOND_JACOBI)
    rep = reps[0]
    print(f"[mg] {which}: {rep}  jacobi iterations {jac[0]['iterations']}", flush=True)
    assert rep["precond"] == abi.PRECOND_MULTIGRID and rep["mg_levels"] >= 2
    assert 0 < rep["mg_coarsest"] <= 1024
    assert rep["converged"] == 1 and rep["rel_residual"] <= 1e-6
    assert jac[0]["precond"] == abi.PRECOND_JACOBI and jac[0]["mg_levels"] == 0
    assert rep["iterations"] < jac[0]["iterations"]
    gt, _ = _gpu(P, nrm, mk(1e-10), dt)
    check_mixed(f"mg {which}", {f: g.field(f) for f in conv}, conv, at_tol, {f: gt.field(f) for f in conv})


@pytest.mark.parametrize("which", ["box", "canyon"])
def test_multigrid_precision_floor(which):
    """Driven to a tight tolerance, the multigrid-preconditioned MIXED
    solve stays within 1e-5 relative L2 of the converged FP64 oracle."""
    P, nrm, mk, dt = _case(which)
    conv = _oracle(P, nrm, mk(1e-12), dt)
    g, reps = _gpu(P, nrm, mk(1e-9), dt)
    assert reps[0]["precond"] == abi.PRECOND_MULTIGRID and reps[0]["converged"] == 1
    for f in conv:
        e = rel_l2(g.field(f), conv[f])
        print(f"[mg-floor] {which} {f}: {e:.2e} ({reps[0]})", flush=True)
        assert e <= 1e-5, f


def test_multigrid_steps_reproducible_graphs_and_direct():
    """Three steps (the hierarchy built at the first, reused after): every
    step converges with the multigrid, and two fresh runs, one with HIP
    graphs and one launching each kernel directly, agree bit for bit."""
    P, nrm, mk, dt = _case("canyon")
    a, ra = _gpu(P, nrm, mk(1e-6), dt, steps=3, graphs=True)
    b, rb = _gpu(P, nrm, mk(1e-6), dt, steps=3, graphs=False)
    for r in ra + rb:
        assert r["precond"] == abi.PRECOND_MULTIGRID and r["converged"] == 1, r
    assert [r["iterations"] for r in ra] == [r["iterations"] for r in rb]
    for f in ("PRES", "UX", "UY", "UZ"):
        np.testing.assert_array_equal(a.field(f), b.field(f), err_msg=f)


def test_multigrid_falls_back_to_jacobi():
    """A PPE of at most 1024 unknowns runs the Jacobi form even when the
    multigrid is asked for; in a batch each case takes its own
    preconditioner."""
    P, nrm = cases.box(nf=6, walls=3)
    prm = cases.box_params(urbansph.default_params(), tol=1e-6)
    _, reps = _gpu(P, nrm, prm, 1e-4)
    assert reps[0]["unknowns"] <= 1024
    assert reps[0]["precond"] == abi.PRECOND_JACOBI and reps[0]["converged"] == 1

    P1, n1, mk, dt = _case("canyon")
    P2, n2 = cases.box(nf=6, walls=3)
    p1 = mk(1e-6)
    p2 = cases.box_params(urbansph.default_params(), tol=1e-6)
    cs = []
    for P, nrm, prm in ((P1, n1, p1), (P2, n2, p2)):
        prm.update(precision=abi.MIXED, ppe_precond=abi.PRECOND_MULTIGRID)
        c = urbansph.Case(prm, len(P))
        c.upload(P, None, nrm)
        cs.append(c)
    b = urbansph.Batch()
    b.step(cs, [0, 0], [dt, dt])
    b.sync()
    r0, r1 = cs[0].ppe_report(), cs[1].ppe_report()
    assert r0.precond == abi.PRECOND_MULTIGRID and r0.mg_levels >= 2 and r0.converged == 1
    assert r1.precond == abi.PRECOND_JACOBI and r1.mg_levels == 0 and r1.converged == 1


def test_multigrid_profiled_build_and_ale():
    """With the per-kernel profile on, a step that builds the hierarchy
    (every step in Lagrangian/ALE mode) records each set-up launch and the
    V-cycle passes, and the step completes."""
    P, nrm, mk, dt = _case("canyon")
    P = P.copy()
    P["eli"] = np.where(P["p_type"] == 1, 1.0, 0.0).astype(np.float32)
    prm = mk(1e-6)
    prm.update(precision=abi.MIXED, ppe_precond=abi.PRECOND_MULTIGRID, ale_positions=1)
    print(f"[mg] profiled ALE n={len(P)}", flush=True)
    g = urbansph.Case(prm, len(P))
    g.upload(P, None, nrm)
    g.profile(True)
    for c in range(2):
        g.step(c, dt)
        g.sync()
        r = g.ppe_report()
        assert r.precond == abi.PRECOND_MULTIGRID and r.converged == 1
    prof = g.profile_report()
    g.profile(False)
    for k in ("mg_rap_terms", "mg_rap_fill", "mg_rap_pack", "mg_gauss_jordan", "mg_res0", "mg_post0", "spmv"):
        assert k in prof and prof[k][0] > 0, k


def _gpu64(P, nrm, prm, dt, steps=1, graphs=True):
    prm.update(precision=abi.FP64, ppe_precond=abi.PRECOND_MULTIGRID)
    g = urbansph.Case(prm, len(P))
    g.options(graphs=graphs, autotune=True)
    g.upload(P, None, nrm)
    for c in range(steps):
        g.step(c, dt)
    g.sync()
    r = g.ppe_report()
    out = {f: g.field(f) for f in ("PRES", "UX", "UY", "UZ")}
    g.close()
    return out, r


@pytest.mark.parametrize("which", ["box", "canyon"])
def test_fp64_multigrid(which):
    """FP64 BiCGSTAB preconditioned by the float V-cycle: the same system in
    double, so at PPE tolerance tol it is no farther from the converged
    oracle than the reference's Jacobi solve at tol (2x, and 1e-9 for the
    rounding of double sums), and driven to 1e-10 it matches the converged
    oracle to 1e-8; in fewer iterations than Jacobi."""
    P, nrm, mk, dt = _case(which)
    conv = _oracle(P, nrm, mk(1e-12), dt)
    at_tol = _oracle(P, nrm, mk(1e-6), dt)
    got, r = _gpu64(P, nrm, mk(1e-6), dt)
    jac = urbansph.Case(mk(1e-6), len(P))
    jac.upload(P, None, nrm)
    jac.step(0, dt)
    jac.sync()
    rj = jac.ppe_report()
    jac.close()
    print(f"[fp64-mg] {which}: iterations {r.iterations} (jacobi {rj.iterations}), levels {r.mg_levels}, "
          f"coarsest {r.mg_coarsest}, rel res {r.rel_residual:.2e}", flush=True)
    assert r.precond == abi.PRECOND_MULTIGRID and r.mg_levels >= 2 and r.converged == 1
    assert r.rel_residual <= 1e-6 and r.iterations < rj.iterations
    e, e64 = mixed_errors(got, conv, at_tol)
    for f in e:
        print(f"[fp64-mg] {which} {f}: multigrid {e[f]:.2e}  jacobi {e64[f]:.2e}", flush=True)
        assert e[f] <= 2.0 * e64[f] + 1e-9, f
    tight, rt = _gpu64(P, nrm, mk(1e-10), dt)
    assert rt.converged == 1 and rt.precond == abi.PRECOND_MULTIGRID
    et, _ = mixed_errors(tight, conv, at_tol)
    print(f"[fp64-mg] {which} at 1e-10: {et}", flush=True)
    assert max(et.values()) <= 1e-8


def test_fp64_multigrid_batch_equals_alone_and_graphs():
    """Two FP64 multigrid cases and a Jacobi one in one batch end bitwise
    where they end alone; graphs on and off give the same bits."""
    specs = []
    for k, nf in enumerate((7, 12, 14)):
        P, nrm = cases.box(nf=nf, walls=3, seed_pos=30 + k, seed_vel=40 + k)
        prm = cases.box_params(urbansph.default_params(), tol=1e-8)
        prm.update(precision=abi.FP64, ppe_precond=abi.PRECOND_MULTIGRID, ppe_solver=abi.PPE_BICGSTAB)
        specs.append((P, nrm, prm, 1e-4))

    def alone(spec, graphs=True):
        P, nrm, prm, dt = spec
        g = urbansph.Case(prm, len(P))
        g.options(graphs=graphs, autotune=True)
        g.upload(P, None, nrm)
        for c in range(3):
            g.step(c, dt)
        g.sync()
        r = g.ppe_report()
        out = {f: g.field(f) for f in ("PRES", "UX", "UY", "UZ")}, (r.iterations, r.precond)
        g.close()
        return out

    solo = [alone(s) for s in specs]
    assert [s[1][1] for s in solo] == [abi.PRECOND_JACOBI, abi.PRECOND_MULTIGRID, abi.PRECOND_MULTIGRID]
    nog = alone(specs[2], graphs=False)
    assert nog[1] == solo[2][1]
    for f in nog[0]:
        np.testing.assert_array_equal(nog[0][f], solo[2][0][f], err_msg=f"graphs {f}")
    b = urbansph.Batch()
    cs = []
    for P, nrm, prm, dt in specs:
        c = urbansph.Case(prm, len(P))
        c.upload(P, None, nrm)
        cs.append(c)
    for st in range(3):
        b.step(cs, [st] * 3, [1e-4] * 3)
    b.sync()
    for k, c in enumerate(cs):
        r = c.ppe_report()
        assert (r.iterations, r.precond) == solo[k][1], k
        for f in solo[k][0]:
            np.testing.assert_array_equal(c.field(f), solo[k][0][f], err_msg=f"case {k} {f}")
        c.close()
    b.close()
