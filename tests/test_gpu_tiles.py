"""Cell-tile record passes (csrc/tiles.hip) against the per-particle passes.

With MIXED precision and float-exact positions, prep, advection, the
pressure force and the PPE assembly can run a workgroup per chunk of a block
of cells, with the chunk's neighbour records staged in LDS and the lists
rewritten to halo slots.  The operands and their order are those of the
per-particle passes, so every field must come out bitwise the same: with
4 x 4 x 2 cell tiles, with shapes whose chunks split (blocks of more particles than
a workgroup) or hold single cells, with a halo cap that sends every chunk to
memory gathers, in a batch of uneven cases, and over several steps with
graph replay.
"""
import numpy as np
import pytest

import abi
import cases
import urbansph
from test_gpu_batch import FIELDS, _assert_same

pytestmark = pytest.mark.gpu

TILE_FIELDS = FIELDS + ["FTOTALY", "FPY", "FLT_S", "VIS_T", "SR"]


TILES = dict(tile_cells=4 | 4 << 8 | 2 << 16, tile_particles=128, tile_halo_max=1536)


def _canyon(n=6e4, **kw):
    P, nrm, geom = cases.sized_canyon(n)
    prm = cases.canyon_params(geom, urbansph.default_params(), tol=1e-6, precision=abi.MIXED)
    prm.update(ppe_precond=abi.PRECOND_MULTIGRID, **kw)
    return P, nrm, prm


def _run(P, nrm, prm, steps=3, dt=5e-5, graphs=True):
    c = urbansph.Case(prm, len(P))
    c.options(graphs=graphs, autotune=True)
    c.upload(P, None, nrm)
    for s in range(steps):
        c.step(s, dt)
    c.sync()
    r = c.ppe_report()
    out = ({f: c.field(f) for f in TILE_FIELDS}, (r.iterations, r.refinements, r.converged, r.rel_residual))
    tiles = (r.tile_chunks, r.tile_raw_chunks, r.tile_halo_largest, r.tile_halo_total)
    c.close()
    return out, tiles


def _same(a, b, what):
    fa, ra = a
    fb, rb = b
    assert ra == rb, f"{what}: PPE report {ra} != {rb}"
    for f in TILE_FIELDS:
        np.testing.assert_array_equal(fa[f], fb[f], err_msg=f"{what}: {f}")


def test_tiles_equal_per_particle():
    P, nrm, prm = _canyon()
    assert prm.tile_cells == 0   # off by default
    ref, t0 = _run(P, nrm, prm)
    assert t0[0] == 0
    got, t = _run(P, nrm, type(prm).from_buffer_copy(prm).update(**TILES))
    print(f"[tiles] n={len(P)} chunks {t[0]} over the cap {t[1]} largest halo {t[2]} "
          f"staged {t[3]} ({t[3] / len(P):.2f} per particle)", flush=True)
    assert t[0] > 0 and t[1] < t[0]
    _same(got, ref, "tiles 4 x 4 x 2")


@pytest.mark.parametrize("cells,width,hmax", [
    (1 | 1 << 8 | 1 << 16, 64, 2048),     # single cells
    (3 | 3 << 8 | 3 << 16, 64, 2048),     # blocks split over several chunks
    (8 | 4 << 8 | 4 << 16, 256, 2048),    # large blocks, wide chunks
    (4 | 4 << 8 | 2 << 16, 128, 64),      # every halo above the cap: memory gathers
])
def test_tile_shapes_equal_per_particle(cells, width, hmax):
    P, nrm, prm = _canyon(3e4)
    ref, _ = _run(P, nrm, type(prm).from_buffer_copy(prm).update(tile_cells=0), steps=2)
    got, t = _run(P, nrm, type(prm).from_buffer_copy(prm).update(tile_cells=cells, tile_particles=width,
                                                                 tile_halo_max=hmax), steps=2)
    print(f"[tiles] cells {cells:#x} width {width} cap {hmax}: chunks {t[0]} over the cap {t[1]} "
          f"largest halo {t[2]}", flush=True)
    assert t[0] > 0
    if hmax == 64:
        assert t[1] == t[0]
    _same(got, ref, f"tiles {cells:#x}/{width}/{hmax}")


def test_tiles_batch_equals_alone():
    specs = []
    for k, (P, nrm, geom, info) in enumerate(cases.ensemble(n_cases=3, lo=8e3, hi=4e4, seed=5)):
        prm = cases.canyon_params(geom, urbansph.default_params(), tol=1e-6, precision=abi.MIXED)
        prm.update(ppe_precond=abi.PRECOND_MULTIGRID, **TILES)
        specs.append((P, nrm, prm, 5e-5 * (1 + 0.1 * k)))
    steps = 3
    alone = []
    for P, nrm, prm, dt in specs:
        out, t = _run(P, nrm, prm, steps=steps, dt=dt)
        assert t[0] > 0
        alone.append(out)
    b = urbansph.Batch()
    cs = []
    for P, nrm, prm, dt in specs:
        c = urbansph.Case(prm, len(P))
        c.upload(P, None, nrm)
        cs.append(c)
    for st in range(steps):
        b.step(cs, [st] * len(cs), [s[3] for s in specs])
    b.sync()
    for k, c in enumerate(cs):
        r = c.ppe_report()
        assert r.tile_chunks > 0
        got = ({f: c.field(f) for f in TILE_FIELDS}, (r.iterations, r.refinements, r.converged, r.rel_residual))
        _same(got, alone[k], f"batch case {k}")
        c.close()
    b.close()
    print(f"[tiles] batch of {len(specs)} equals alone", flush=True)
