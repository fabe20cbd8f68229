"""Stencil slices of the MIXED PPE against slices that store their columns.

With fixed particles on a lattice and the rows in lattice order, a 64-row
slice whose rows' column offsets (col - row) have a small union stores every
row over that union, an offset a row lacks holding the value 0, and its
products read no columns (ppe_stencil = 1, the default); after the first
assembly of a search the slices drop the offsets whose values are zero in
all their rows.  The nonzero entries keep their ascending column order and
the stored zeros add exact zeros, so every field and the PPE report come out
bitwise the same as with
ppe_stencil = 0: on the canyon with the multigrid and with Jacobi over
several steps (graph replay), on an APR-refined canyon, and in a batch.
"""
import numpy as np
import pytest

import abi
import cases
import urbansph
from test_gpu_apr import _refined
from test_gpu_tiles import TILE_FIELDS, _same

pytestmark = pytest.mark.gpu


def _run(P, nrm, prm, steps=4, dt=5e-5):
    c = urbansph.Case(prm, len(P))
    c.options(graphs=True, autotune=True)
    c.upload(P, None, nrm)
    for s in range(steps):
        c.step(s, dt)
    c.sync()
    r = c.ppe_report()
    out = ({f: c.field(f) for f in TILE_FIELDS}, (r.iterations, r.refinements, r.converged, r.rel_residual))
    info = dict(slices=r.stencil_slices, nnz_st=r.stencil_nnz, nnz=r.nnz, slots=r.slots, unknowns=r.unknowns,
                ordering=r.ordering, precond=r.precond)
    c.close()
    return out, info


@pytest.mark.parametrize("precond", [abi.PRECOND_MULTIGRID, abi.PRECOND_JACOBI])
def test_stencil_slices_equal_stored_columns(precond):
    P, nrm, geom = cases.sized_canyon(6e4)
    prm = cases.canyon_params(geom, urbansph.default_params(), tol=1e-6, precision=abi.MIXED)
    prm.update(ppe_precond=precond)
    assert prm.ppe_stencil == 1   # the default
    got, st = _run(P, nrm, prm)
    ref, ex = _run(P, nrm, type(prm).from_buffer_copy(prm).update(ppe_stencil=0))
    nslices = (st["unknowns"] + 63) // 64
    print(f"[stencil] n={len(P)} rows {st['unknowns']}: stencil slices {st['slices']} of {nslices}, "
          f"entries {st['nnz_st']} of {st['nnz']}, slots {st['slots']} (columns stored: {ex['slots']})", flush=True)
    assert st["ordering"] == abi.ORDER_LATTICE
    assert ex["slices"] == 0 and ex["nnz_st"] == 0
    assert st["slices"] > nslices // 2 and 0 < st["nnz_st"] <= st["nnz"]
    # after the first assembly the stencil slices drop the offsets whose
    # values are zero (the search-radius neighbours); the rest of the union's
    # slots are stored zeros
    assert st["nnz"] < ex["nnz"]
    assert st["slots"] <= 5 * ex["slots"] // 4 + 64
    _same(got, ref, f"stencil slices, precond {precond}")


def test_stencil_slices_refined():
    _, _, geom = cases.small_canyon(dx=0.02)
    prm = cases.canyon_params(geom, urbansph.default_params(), tol=1e-6, precision=abi.MIXED)
    prm.update(ppe_precond=abi.PRECOND_MULTIGRID)
    got, live, r, n, n0 = _refined(prm)
    ref, live_r, r_r, n_r, _ = _refined(type(prm).from_buffer_copy(prm).update(ppe_stencil=0))
    print(f"[stencil] refined n {n0} -> {n}: ordering {r.ordering}, stencil slices {r.stencil_slices}, "
          f"entries {r.stencil_nnz} of {r.nnz}", flush=True)
    assert r_r.stencil_slices == 0
    assert n == n_r and np.array_equal(live, live_r)
    assert (r.iterations, r.refinements, r.rel_residual) == (r_r.iterations, r_r.refinements, r_r.rel_residual)
    for f in got:
        np.testing.assert_array_equal(got[f], ref[f], err_msg=f)


def test_stencil_slices_batch():
    specs = []
    for k, (P, nrm, geom, info) in enumerate(cases.ensemble(n_cases=3, lo=8e3, hi=4e4, seed=11)):
        prm = cases.canyon_params(geom, urbansph.default_params(), tol=1e-6, precision=abi.MIXED)
        prm.update(ppe_precond=abi.PRECOND_MULTIGRID)
        specs.append((P, nrm, prm, 5e-5 * (1 + 0.1 * k)))
    steps = 3
    outs = {}
    for flag in (1, 0):
        b = urbansph.Batch()
        cs = []
        for P, nrm, prm, dt in specs:
            c = urbansph.Case(type(prm).from_buffer_copy(prm).update(ppe_stencil=flag), len(P))
            c.upload(P, None, nrm)
            cs.append(c)
        for st in range(steps):
            b.step(cs, [st] * len(cs), [s[3] for s in specs])
        b.sync()
        res = []
        for c in cs:
            r = c.ppe_report()
            assert (r.stencil_slices > 0) == (flag == 1)
            res.append(({f: c.field(f) for f in TILE_FIELDS}, (r.iterations, r.refinements, r.converged,
                                                                r.rel_residual)))
            c.close()
        b.close()
        outs[flag] = res
    for k in range(len(specs)):
        _same(outs[1][k], outs[0][k], f"batch case {k}")
    print(f"[stencil] batch of {len(specs)}: stencil slices equal stored columns", flush=True)


def test_stencil_not_built_for_moving_particles():
    P, nrm, geom = cases.sized_canyon(2e4)
    P = P.copy()
    P["eli"] = np.where(P["p_type"] == 1, 1.0, 0.0).astype(np.float32)   # the fluid moves
    prm = cases.canyon_params(geom, urbansph.default_params(), tol=1e-6, precision=abi.MIXED)
    prm.update(ppe_precond=abi.PRECOND_MULTIGRID, ale_positions=1)
    _, info = _run(P, nrm, prm, steps=2)
    assert info["slices"] == 0 and info["nnz_st"] == 0
