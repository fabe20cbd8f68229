"""Adaptive particle refinement through the C ABI against the oracle:
the same particles split in every pass, the searches after refinement are
bit-exact, and the refined case steps to the FP64 tolerance."""
import numpy as np
import pytest

import abi
import cases
import oracle
import urbansph
from test_gpu_parity import check_mixed

pytestmark = pytest.mark.gpu

FIELDS = ["UX", "UY", "UZ", "PRES", "VIS_T", "FTOTALX", "FPX", "FPZ", "FBX", "INV_CM_XX", "X", "Y", "Z", "M", "H"]


def rel_l2(a, b):
    s = np.linalg.norm(b)
    return np.linalg.norm(a - b) / s if s > 0 else np.linalg.norm(a - b)


def _targets(case, geom_buildings, r1, r2):
    x, y, z = case.field("X"), case.field("Y"), case.field("Z")
    return cases.apr_targets(x, y, z, case.ifield("P_TYPE"), case.ifield("I_TYPE"), geom_buildings, r1, r2)


@pytest.mark.parametrize("passes", [1, 40])
def test_apr_canyon_fp64(passes):
    P, nrm, geom = cases.small_canyon(dx=0.02)
    prm = cases.canyon_params(geom, urbansph.default_params(), tol=1e-10)
    cap = 40 * len(P)
    o = oracle.Oracle(prm, cap)
    o.upload(P, None, nrm)
    g = urbansph.Case(prm, cap)
    g.upload(P, None, nrm)
    dt = 5e-5
    assert o.step(0, dt) == 0
    g.step(0, dt)
    t = _targets(g, geom["buildings"], 0.05, 0.03)
    print(f"[apr] n={len(P)} targets: {int((t == 1).sum())} at level 1, {int((t == 2).sum())} at level 2", flush=True)
    n_o = o.apr_refine(t, max_passes=passes)
    n_g = g.apr_refine(t, max_passes=passes)
    print(f"[apr] passes={passes}: n {len(P)} -> {n_g}", flush=True)
    assert n_g == n_o and n_g > len(P)
    ko, po, so, eo = o.nnps()
    kg, pg, sg, eg = g.nnps()
    np.testing.assert_array_equal(kg, ko)
    np.testing.assert_array_equal(pg, po)
    np.testing.assert_array_equal(sg, so)
    np.testing.assert_array_equal(eg, eo)
    for kind in (0, 1):
        oo, no = o.neighbors(kind)
        og, ng = g.neighbors(kind)
        np.testing.assert_array_equal(og, oo)
        np.testing.assert_array_equal(ng, no)
    for c in (1, 2):
        assert o.step(c, dt) == 0
        g.step(c, dt)
    g.sync()
    live = g.ifield("I_TYPE") != 3
    for f in FIELDS:
        a, b = g.field(f)[live], o.field(f)[live]
        assert np.all(np.isfinite(a)), f
        e = rel_l2(a, b)
        assert e <= 1e-10, f"{f}: {e:.3e}"


def _refined(prm, steps=3, graphs=True):
    P, nrm, geom = cases.small_canyon(dx=0.02)
    g = urbansph.Case(prm, 40 * len(P))
    g.options(graphs=graphs, autotune=True)
    g.upload(P, None, nrm)
    g.step(0, 5e-5)
    n = g.apr_refine(_targets(g, geom["buildings"], 0.05, 0.03), max_passes=40)
    for c in range(1, 1 + steps):
        g.step(c, 5e-5)
    g.sync()
    out = {f: g.field(f) for f in ("UX", "UY", "UZ", "PRES")}
    live = g.ifield("I_TYPE") != 3
    r = g.ppe_report()
    g.close()
    return out, live, r, n, len(P)


def _refined_oracle(prm, steps=3):
    """The refined case of _refined on the oracle (FP64)."""
    P, nrm, geom = cases.small_canyon(dx=0.02)
    o = oracle.Oracle(prm, 40 * len(P))
    o.upload(P, None, nrm)
    assert o.step(0, 5e-5) == 0
    x, y, z = o.field("X"), o.field("Y"), o.field("Z")
    t = cases.apr_targets(x, y, z, o.ifield("P_TYPE"), o.ifield("I_TYPE"), geom["buildings"], 0.05, 0.03)
    o.apr_refine(t, max_passes=40)
    for c in range(1, 1 + steps):
        assert o.step(c, 5e-5) == 0
    return {f: o.field(f) for f in ("UX", "UY", "UZ", "PRES")}


def test_apr_mixed_runs():
    """MIXED on a refined case: rows of more than 256 entries (coarse rows
    beside refined particles) keep the rest of their entries in a contiguous
    tail summed by a wave each; the result stays as close to a tightly
    converged FP64 solve as the PPE tolerance allows."""
    _, _, geom = cases.small_canyon(dx=0.02)
    prm = cases.canyon_params(geom, urbansph.default_params(), tol=1e-6, precision=abi.MIXED)
    tight = type(prm).from_buffer_copy(prm)
    tight.update(precision=abi.FP64, ppe_tol=1e-11)
    at_tol = type(prm).from_buffer_copy(prm)
    at_tol.update(precision=abi.FP64)
    out, live, r, n, n0 = _refined(prm)
    conv, tol6 = _refined_oracle(tight), _refined_oracle(at_tol)
    assert r.converged == 1 and r.rel_residual <= 1e-6
    assert r.tail_rows > 0
    assert n > n0
    for f in out:
        assert np.all(np.isfinite(out[f][live])), f
    sel = lambda d: {f: d[f][live] for f in d}
    tight = type(prm).from_buffer_copy(prm)
    tight.update(ppe_tol=1e-10)
    floor, live_t, _, _, _ = _refined(tight)
    assert np.array_equal(live_t, live)
    check_mixed("apr mixed", sel(out), sel(conv), sel(tol6), sel(floor))


def test_apr_multigrid_runs():
    """The multigrid preconditioner on a refined case: aggregation on each
    particle's own lattice, long rows (tails) in the level-0 passes and
    in the Galerkin products; every solve converges with it, in fewer
    iterations than the Jacobi form, and the result stays as close to a
    tightly converged FP64 solve as the PPE tolerance allows."""
    _, _, geom = cases.small_canyon(dx=0.02)
    prm = cases.canyon_params(geom, urbansph.default_params(), tol=1e-6, precision=abi.MIXED)
    prm.update(ppe_precond=abi.PRECOND_MULTIGRID)
    jac = type(prm).from_buffer_copy(prm)
    jac.update(ppe_precond=abi.PRECOND_JACOBI)
    tight = type(prm).from_buffer_copy(prm)
    tight.update(precision=abi.FP64, ppe_tol=1e-11, ppe_precond=abi.PRECOND_JACOBI)
    at_tol = type(prm).from_buffer_copy(prm)
    at_tol.update(precision=abi.FP64, ppe_precond=abi.PRECOND_JACOBI)
    out, live, r, n, n0 = _refined(prm)
    _, _, rj, _, _ = _refined(jac)
    conv, tol6 = _refined_oracle(tight), _refined_oracle(at_tol)
    print(f"[apr mg] levels {r.mg_levels} coarsest {r.mg_coarsest} iterations {r.iterations} "
          f"(jacobi {rj.iterations}) tails {r.tail_rows}", flush=True)
    assert r.precond == abi.PRECOND_MULTIGRID and r.mg_levels >= 2
    assert r.converged == 1 and r.rel_residual <= 1e-6
    assert r.tail_rows > 0 and n > n0
    assert r.iterations < rj.iterations
    for f in out:
        assert np.all(np.isfinite(out[f][live])), f
    sel = lambda d: {f: d[f][live] for f in d}
    tight = type(prm).from_buffer_copy(prm)
    tight.update(ppe_tol=1e-10)
    floor, live_t, _, _, _ = _refined(tight)
    assert np.array_equal(live_t, live)
    check_mixed("apr mg", sel(out), sel(conv), sel(tol6), sel(floor))


def test_apr_needs_room_and_a_search():
    P, _ = cases.box(nf=6, walls=3)
    prm = cases.box_params(urbansph.default_params())
    g = urbansph.Case(prm, len(P) + 8)
    g.upload(P)
    with pytest.raises(urbansph.SPHError) as e:
        g.apr_refine(np.ones(len(P)), max_passes=1)
    assert e.value.rc == abi.E_STATE
    g.step(0, 1e-4)
    with pytest.raises(urbansph.SPHError) as e:
        g.apr_refine(np.ones(len(P)), max_passes=1)
    assert e.value.rc == abi.E_NOMEM



def test_refined_mixed_case_keeps_cell_order():
    """A refined case has rows longer than the lattice-order sort holds:
    MIXED keeps cell order for it, and its 16-bit columns, whose windows are
    collected in any order, still encode the same matrix (same iteration
    count and result as 32-bit columns)."""
    P, nrm, geom = cases.small_canyon(dx=0.02)
    out = []
    for i16 in (1, 0):
        prm = cases.canyon_params(geom, urbansph.default_params(), tol=1e-6, precision=abi.MIXED)
        prm.update(ppe_index16=i16)
        g = urbansph.Case(prm, 40 * len(P))
        g.upload(P, None, nrm)
        g.step(0, 5e-5)
        g.apr_refine(_targets(g, geom["buildings"], 0.05, 0.03), max_passes=40)
        g.step(1, 5e-5)
        g.sync()
        r = g.ppe_report()
        out.append((r.ordering, r.index16, r.iterations, g.field("PRES")))
        g.close()
    assert out[0][0] == abi.ORDER_CELL and out[0][1] == 1 and out[1][1] == 0
    assert out[0][2] == out[1][2]
    np.testing.assert_array_equal(out[0][3], out[1][3])


def test_apr_multigrid_reproducible_graphs_and_direct():
    """The refined case on the multigrid (hashed and sorted Galerkin rows,
    wave-per-row products of long coarse rows, long neighbour lists last in
    the record-gathering passes): two fresh runs, one with HIP graphs and one
    launching each kernel directly, agree bit for bit."""
    _, _, geom = cases.small_canyon(dx=0.02)
    prm = cases.canyon_params(geom, urbansph.default_params(), tol=1e-6, precision=abi.MIXED)
    prm.update(ppe_precond=abi.PRECOND_MULTIGRID)
    print("[apr mg repro] graphs / direct", flush=True)
    a, _, ra, _, _ = _refined(prm, graphs=True)
    b, _, rb, _, _ = _refined(prm, graphs=False)
    print(f"[apr mg repro] sorted rows {ra.mg_sorted_rows}, long coarse rows {ra.mg_long_rows}, "
          f"heavy lists {ra.heavy_lists}, tails {ra.tail_rows}", flush=True)
    assert ra.precond == abi.PRECOND_MULTIGRID and rb.precond == abi.PRECOND_MULTIGRID
    assert ra.iterations == rb.iterations
    # the paths the docstring names all ran
    assert ra.tail_rows > 0 and ra.mg_sorted_rows > 0 and ra.mg_long_rows > 0 and ra.heavy_lists > 0
    for f in a:
        np.testing.assert_array_equal(a[f], b[f], err_msg=f)


def test_sorting_galerkin_kernels_agree_with_the_hash():
    """Every coarse row of the refined case formed by the sorting kernels
    (double sums in term order) instead of the fixed-point hash: the coarse
    operators agree to float rounding, so the solve takes the same number of
    iterations (one more at most) and ends within the tolerance of the
    hashed hierarchy's result."""
    _, _, geom = cases.small_canyon(dx=0.02)
    prm = cases.canyon_params(geom, urbansph.default_params(), tol=1e-6, precision=abi.MIXED)
    prm.update(ppe_precond=abi.PRECOND_MULTIGRID)
    srt = type(prm).from_buffer_copy(prm)
    srt.update(mg_sort_all=1)
    a, live, ra, _, _ = _refined(prm)
    b, _, rb, _, _ = _refined(srt)
    print(f"[apr sort-all] iterations hash {ra.iterations} sort {rb.iterations}; sorted rows {ra.mg_sorted_rows} / "
          f"{rb.mg_sorted_rows}", flush=True)
    assert rb.precond == abi.PRECOND_MULTIGRID and rb.converged == 1
    assert rb.mg_sorted_rows > ra.mg_sorted_rows and rb.mg_levels == ra.mg_levels
    assert abs(rb.iterations - ra.iterations) <= 1
    for f in a:
        e = rel_l2(b[f][live], a[f][live])
        print(f"[apr sort-all] {f}: {e:.2e}", flush=True)
        assert e <= 1e-5, f
